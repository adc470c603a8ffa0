"""Synthetic initial conditions for the SPH derivative pass.

The lattice follows GenerateNodeDistribution3d.latticeDistribution
(src/NodeGenerators/GenerateNodeDistribution3d.py:573-626): index
i + nx*(j + ny*k), cell-centred positions xmin + (i + 1/2) dx, mass
dx*dy*dz*rho, H = diag(1/(nPerh dx)).  Jitter (seeded, uniform in +-jitter*dx
per coordinate) is added on top, as the benchmark configurations ask.
"""
import numpy as np


def sym_identity(dim, n, value):
    """n SymTensors value*I in storage order (xx,xy,xz,yy,yz,zz)."""
    ns = dim * (dim + 1) // 2
    H = np.zeros((n, ns))
    k = 0
    for a in range(dim):
        for b in range(a, dim):
            if a == b:
                H[:, k] = value
            k += 1
    return H


def lattice(dim, n, jitter=0.0, seed=0, nPerh=1.51, xmin=0.0, xmax=1.0, rho=1.0):
    """n^dim nodes in [xmin, xmax]^dim.  Returns pos (N,dim), mass (N,), H (N,nsym)."""
    N = n ** dim
    dx = (xmax - xmin) / n
    idx = np.arange(N, dtype=np.int64)
    pos = np.zeros((N, dim))
    rem = idx.copy()
    for a in range(dim):
        ia = rem % n
        rem //= n
        pos[:, a] = xmin + (ia + 0.5) * dx
    if jitter:
        rng = np.random.default_rng(seed)
        pos += rng.uniform(-jitter, jitter, size=(N, dim)) * dx
    vol = 1.0
    for _ in range(dim):
        vol = vol * dx
    rho_arr = np.full(N, rho) if np.isscalar(rho) else rho(pos)
    mass = vol * rho_arr
    H = sym_identity(dim, N, 1.0 / (nPerh * dx))
    return pos, mass, H, dx


def gamma_law(rho, eps, gamma):
    """GammaLawGas pressure and sound speed (src/Material/GammaLawGas.cc:163-170, 215-223),
    no pressure limits active."""
    g1 = gamma - 1.0
    P = (g1 * rho) * eps
    cs = np.sqrt(np.maximum(0.0, (gamma * g1) * eps))
    return P, cs


def converging_flow(n, dim=3, jitter=0.01, seed=12345, nPerh=1.51, eps0=1.0e-6, gamma=5.0 / 3.0):
    """Cold, uniform-density gas with radial inflow v = -r_hat about the box centre
    (3-D converging-flow configuration; 32^3 and 128^3 in the benchmarks)."""
    pos, mass, H, dx = lattice(dim, n, jitter=jitter, seed=seed, nPerh=nPerh)
    N = pos.shape[0]
    rho = np.ones(N)
    eps = np.full(N, eps0)
    d = pos - 0.5
    r = np.sqrt((d * d).sum(axis=1))
    vel = -d / np.maximum(r, 1e-300)[:, None]
    P, cs = gamma_law(rho, eps, gamma)
    return dict(pos=pos, vel=vel, mass=mass, rho=rho, eps=eps, P=P, cs=cs, H=H,
                omega=np.ones(N)), dict(n_internal=N, dx=dx, gamma=gamma, nPerh=nPerh)


def random_state(dim, n, seed=7, jitter=0.2, nPerh=1.51, gamma=1.4, vel_scale=0.3, rho_var=0.2,
                 h_var=0.0):
    """Perturbed lattice with random velocity, density and energy fields, for
    parity tests that exercise every branch of the pair loop (compression and
    expansion, Q on and off).  h_var > 0 gives each node its own h."""
    pos, mass, H, dx = lattice(dim, n, jitter=jitter, seed=seed, nPerh=nPerh)
    N = pos.shape[0]
    rng = np.random.default_rng(seed + 1)
    rho = 1.0 + rho_var * rng.uniform(-1, 1, N)
    mass = mass * rho
    eps = 1.0 + 0.5 * rng.uniform(-1, 1, N)
    vel = vel_scale * rng.standard_normal((N, dim))
    if h_var:
        scale = 1.0 + h_var * rng.uniform(-1, 1, N)
        H = H * scale[:, None]
    P, cs = gamma_law(rho, eps, gamma)
    return dict(pos=pos, vel=vel, mass=mass, rho=rho, eps=eps, P=P, cs=cs, H=H,
                omega=np.ones(N)), dict(n_internal=N, dx=dx, gamma=gamma, nPerh=nPerh)


def anisotropic_H(dim, n, h0, seed=3, spread=0.3):
    """SymTensors R diag(1/h) R^T with random rotations and +-spread anisotropy."""
    rng = np.random.default_rng(seed)
    ns = dim * (dim + 1) // 2
    H = np.zeros((n, ns))
    for i in range(n):
        h = h0 * (1.0 + spread * rng.uniform(-1, 1, dim))
        Q, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
        Mx = Q @ np.diag(1.0 / h) @ Q.T
        k = 0
        for a in range(dim):
            for b in range(a, dim):
                H[i, k] = 0.5 * (Mx[a, b] + Mx[b, a])
                k += 1
    return H


def linear_velocity(dim, n, A, seed=14892042, jitter=0.2, nPerh=1.51):
    """v = A.x on a perturbed lattice (src/SPH/tests/testLinearVelocityGradient.py)."""
    pos, mass, H, dx = lattice(dim, n, jitter=jitter, seed=seed, nPerh=nPerh)
    N = pos.shape[0]
    A = np.asarray(A, dtype=np.float64).reshape(dim, dim)
    vel = pos @ A.T
    rho = np.ones(N)
    eps = np.ones(N)
    P, cs = gamma_law(rho, eps, 5.0 / 3.0)
    return dict(pos=pos, vel=vel, mass=mass, rho=rho, eps=eps, P=P, cs=cs, H=H,
                omega=np.ones(N)), dict(n_internal=N, dx=dx, nPerh=nPerh)


def gaussian_velocity_field(n, dim=3, sigma=0.1, slope=-2.0, seed=2024):
    """Seeded Gaussian random velocity field on the n^dim lattice: white noise
    shaped in Fourier space to a power spectrum ~ k^slope (|v_k| ~ k^(slope/2)),
    zero mean, each component scaled to RMS sigma."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n) * n
    grids = np.meshgrid(*([k1] * dim), indexing="ij")
    kmag = np.sqrt(sum(g * g for g in grids))
    amp = np.zeros_like(kmag)
    nz = kmag > 0
    amp[nz] = kmag[nz] ** (0.5 * slope)
    vel = np.zeros((n ** dim, dim))
    for a in range(dim):
        noise = rng.standard_normal([n] * dim)
        f = np.fft.ifftn(np.fft.fftn(noise) * amp).real
        f -= f.mean()
        f *= sigma / max(f.std(), 1e-300)
        # lattice() numbers nodes x-fastest: index = ix + n*iy + n^2*iz
        vel[:, a] = f.transpose(*reversed(range(dim))).reshape(-1)
    return vel


def periodic_sinusoid(n, dim=3, amp=0.1, sigma=0.1, gamma=1.4, u=1.0, seed=2024, nPerh=1.51):
    """Periodic-box workload: unit lattice, rho = 1 + amp * prod sin(2 pi x_a),
    masses m = rho(x) dx^dim, Gaussian random velocity (sigma, k^-2), gamma-law
    gas with specific energy u."""
    pos, mass, H, dx = lattice(dim, n, nPerh=nPerh)
    N = pos.shape[0]
    rho = 1.0 + amp * np.prod(np.sin(2.0 * np.pi * pos), axis=1)
    mass = rho * dx ** dim
    eps = np.full(N, u)
    vel = gaussian_velocity_field(n, dim, sigma=sigma, seed=seed)
    P, cs = gamma_law(rho, eps, gamma)
    return dict(pos=pos, vel=vel, mass=mass, rho=rho, eps=eps, P=P, cs=cs, H=H,
                omega=np.ones(N)), dict(n_internal=N, dx=dx, gamma=gamma, nPerh=nPerh)


def shock_tube(n, rhoL=1.0, PL=1.0, rhoR=0.125, PR=0.1, gamma=1.4, nPerh=1.51):
    """Planar two-state tube on [0, 1] split at 0.5, equal node masses, the
    nodes on each side in proportion to its mass (the layout of
    tests/Hydro/Sod/Sod-planar-1d.py:14-32, 189-231), v = 0."""
    mtot = 0.5 * rhoL + 0.5 * rhoR
    m = mtot / n
    nL = int(round(n * 0.5 * rhoL / mtot))
    nL = min(max(nL, 1), n - 1)
    nR = n - nL
    xL = (np.arange(nL) + 0.5) * (0.5 / nL)
    xR = 0.5 + (np.arange(nR) + 0.5) * (0.5 / nR)
    pos = np.concatenate([xL, xR])[:, None]
    rho = np.concatenate([np.full(nL, rhoL), np.full(nR, rhoR)])
    P = np.concatenate([np.full(nL, PL), np.full(nR, PR)])
    eps = P / ((gamma - 1.0) * rho)
    dx = np.concatenate([np.full(nL, 0.5 / nL), np.full(nR, 0.5 / nR)])
    H = (1.0 / (nPerh * dx))[:, None]
    return dict(pos=pos, vel=np.zeros((n, 1)), mass=np.full(n, m), rho=rho, eps=eps, H=H)


def tube_ensemble(count=512, nmin=200, nmax=20000, t_end=0.2, seed=4242):
    """Batched shock tubes: node counts log-uniform in [nmin, nmax], left
    rho, P ~ U[0.5, 2], right rho, P ~ U[0.05, 0.5], v = 0, gamma in
    {1.4, 5/3}, each run to t_end.  Returns a list of dicts (state, gamma,
    t_end, n)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).round().astype(np.int64)
    rhoL, PL = rng.uniform(0.5, 2.0, count), rng.uniform(0.5, 2.0, count)
    rhoR, PR = rng.uniform(0.05, 0.5, count), rng.uniform(0.05, 0.5, count)
    gam = np.where(rng.uniform(0.0, 1.0, count) < 0.5, 1.4, 5.0 / 3.0)
    out = []
    for q in range(count):
        st = shock_tube(int(ns[q]), rhoL[q], PL[q], rhoR[q], PR[q], gam[q])
        out.append(dict(state=st, gamma=float(gam[q]), t_end=float(t_end), n=int(ns[q])))
    return out


def bspline_w(dim, eta):
    """Cubic B-spline kernel W(eta) with unit det H (BSplineKernelInline.hh:9-100)."""
    A = {1: 2.0 / 3.0, 2: 10.0 / (7.0 * np.pi), 3: 1.0 / np.pi}[dim]
    eta = np.asarray(eta, dtype=np.float64)
    w = np.where(eta < 1.0, 1.0 - 1.5 * eta ** 2 + 0.75 * eta ** 3,
                 np.where(eta < 2.0, 0.25 * (2.0 - eta) ** 3, 0.0))
    return A * w


def point_blast(n, gamma=5.0 / 3.0, E=1.0, P0=1.0e-6, jitter=0.01, seed=77, spike_scale=0.5, nPerh=1.51):
    """2-D point blast: n x n jittered lattice on [0, 1]^2, rho = 1, P = P0,
    v = 0, energy E deposited with the smoothed-spike rule of
    tests/Hydro/Sedov/Sedov-cylindrical-2d.py:240-262 about the box centre:
    eps_i = eps0 + E W(|H_i (x_i - c)|/spike_scale) / (m_i sum W)."""
    pos, mass, H, dx = lattice(2, n, jitter=jitter, seed=seed, nPerh=nPerh)
    N = pos.shape[0]
    eps0 = P0 / ((gamma - 1.0) * 1.0)
    eta = np.sqrt((((pos - 0.5) * H[:, [0]]) ** 2).sum(1))
    W = bspline_w(2, eta / spike_scale)
    eps = eps0 + E * W / mass / W.sum()
    return dict(pos=pos, vel=np.zeros((N, 2)), mass=mass, rho=np.ones(N), eps=eps, H=H)


def blast_requests(count, rate, nmin=1000, nmax=250000, tmin=0.05, tmax=0.2, seed=99):
    """Poisson arrivals (rate per second) of 2-D point blasts with node counts
    log-uniform in [nmin, nmax] (rounded to squares) and end times U[tmin,
    tmax].  Returns a list of dicts (state, gamma, t_end, arrival, n)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, count)
    arrival = np.cumsum(gaps) - gaps[0]
    side = np.sqrt(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))).round().astype(np.int64)
    tend = rng.uniform(tmin, tmax, count)
    out = []
    for q in range(count):
        st = point_blast(int(side[q]), seed=int(rng.integers(1 << 30)))
        out.append(dict(state=st, gamma=5.0 / 3.0, t_end=float(tend[q]), arrival=float(arrival[q]),
                        n=int(side[q]) ** 2))
    return out
