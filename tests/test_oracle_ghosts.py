"""Oracle checks for planar ghost boundaries (src/Boundary/PlanarBoundary.cc,
ReflectingBoundary.cc, PeriodicBoundary.cc, src/Integrator/Integrator.cc:443-595).

Pinned by closed forms: ghost counts of a lattice in a box, exact mirror and
translation maps for axis-aligned planes, eigenvalues against LAPACK, the
uniform lattice that ghosts make infinite (no boundary forces), and brute-force
culling."""
import numpy as np
import pytest

from sphgpu import problems


def counts_closed_form(dim, n, layers=3):
    """Ghosts of a reflecting box around an n^dim lattice: each axis adds
    2*layers slabs whose cross-section includes the ghosts of earlier axes."""
    total, width = 0, [n] * dim
    for a in range(dim):
        cross = int(np.prod([width[b] for b in range(dim) if b != a]))
        total += 2 * layers * cross
        width[a] += 2 * layers
    return total


@pytest.mark.parametrize("dim,n", [(1, 40), (2, 12), (3, 8)])
@pytest.mark.parametrize("kind", ["reflecting", "periodic"])
def test_ghost_count_closed_form(oracle, dim, n, kind):
    st, meta = problems.converging_flow(n, dim=dim)
    b = oracle.box_boundaries(dim, kind=kind)
    pos, H, control, sub_of, sub_start = oracle.set_ghosts(dim, st["pos"], st["H"], b)
    N = st["pos"].shape[0]
    assert pos.shape[0] - N == counts_closed_form(dim, n)
    assert np.all(np.diff(sub_of) >= 0)
    assert np.all(control < np.arange(N, pos.shape[0]))       # a control precedes its ghost
    assert sub_start[-1] == pos.shape[0] - N


def test_reflection_maps(oracle):
    dim, n = 3, 8
    st, meta = problems.converging_flow(n, dim=dim)
    N = st["pos"].shape[0]
    b = oracle.box_boundaries(dim)
    pos, H, control, sub_of, _ = oracle.set_ghosts(dim, st["pos"], st["H"], b)
    vel = np.zeros((pos.shape[0], dim))
    vel[:N] = st["vel"]
    vel = oracle.apply_ghosts(dim, N, b, control, sub_of, vel, 1)
    g = np.arange(N, pos.shape[0])
    c = control
    for s in range(6):
        m = sub_of == s
        axis, hi = s // 2, s % 2
        other = [a for a in range(dim) if a != axis]
        if hi == 0:     # plane x_axis = 0: exact negation
            assert np.array_equal(pos[g[m], axis], -pos[c[m], axis])
        else:           # plane x_axis = 1: 2 - x to rounding
            assert np.allclose(pos[g[m], axis], 2.0 - pos[c[m], axis], rtol=0, atol=4e-16)
        assert np.array_equal(pos[g[m]][:, other], pos[c[m]][:, other])
        assert np.array_equal(vel[g[m], axis], -vel[c[m], axis])
        assert np.array_equal(vel[g[m]][:, other], vel[c[m]][:, other])
        assert np.array_equal(H[g[m]], H[c[m]] * 1.0)            # isotropic H is invariant (+-0 off-diagonals)


def test_periodic_maps(oracle):
    dim, n = 2, 12
    st, meta = problems.converging_flow(n, dim=dim)
    N = st["pos"].shape[0]
    b = oracle.box_boundaries(dim, kind="periodic")
    pos, H, control, sub_of, _ = oracle.set_ghosts(dim, st["pos"], st["H"], b)
    g = np.arange(N, pos.shape[0])
    for s in range(4):
        m = sub_of == s
        axis = s // 2
        shift = -1.0 if s % 2 == 0 else 1.0     # plane1 -> plane2 maps x ~ 1 to x - 1; then x ~ 0 to x + 1
        assert np.array_equal(pos[g[m], axis], pos[control[m], axis] + shift)


def test_min_eigen_vs_lapack(oracle):
    rng = np.random.default_rng(5)
    for dim in (2, 3):
        Hs = problems.anisotropic_H(dim, 50, 0.05, seed=9)
        for h in Hs:
            M = np.zeros((dim, dim))
            k = 0
            for a in range(dim):
                for c in range(a, dim):
                    M[a, c] = M[c, a] = h[k]
                    k += 1
            assert oracle.min_eigen(dim, h) == pytest.approx(np.linalg.eigvalsh(M)[0], rel=1e-10)
    # isotropic: within a few ulp of the diagonal
    for dim in (1, 2, 3):
        h = problems.sym_identity(dim, 1, 37.75)[0]
        assert oracle.min_eigen(dim, h) == pytest.approx(37.75, rel=1e-14)
    _ = rng


@pytest.mark.parametrize("kind", ["reflecting", "periodic"])
def test_uniform_lattice_has_no_boundary(oracle, kind):
    """Ghosts continue the lattice: every internal node sees a full stencil,
    the same density, and no net force."""
    dim, n = 3, 10
    pos, mass, H, dx = problems.lattice(dim, n)
    N = pos.shape[0]
    rho, eps = np.ones(N), np.full(N, 2.0)
    P, cs = problems.gamma_law(rho, eps, 5 / 3)
    st = dict(pos=pos, vel=np.zeros((N, dim)), mass=mass, rho=rho, eps=eps, P=P, cs=cs, H=H, omega=np.ones(N))
    b = oracle.box_boundaries(dim, kind=kind)
    full, off, nbr, info = oracle.ghost_pipeline(dim, st, b)
    assert np.all(np.diff(off) == 122)
    T = oracle.Table(dim, 1000)
    rhoSum = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr)
    assert np.allclose(rhoSum, rhoSum[0], rtol=1e-13, atol=0)
    out = oracle.eval_derivs(T, oracle.make_params(), N, full, off, nbr)
    # free-surface force of the same lattice without ghosts sets the scale
    off0, nbr0 = oracle.connectivity(dim, N, pos, H)
    out0 = oracle.eval_derivs(T, oracle.make_params(), N, st, off0, nbr0)
    scale = np.abs(out0["DvDt"]).max()
    assert np.abs(out["DvDt"]).max() < 1e-12 * scale
    assert np.all(out["DepsDt"] == 0)
    assert np.all(out["DvDx"] == 0)


@pytest.mark.parametrize("dim,n", [(2, 10), (3, 6)])
def test_culling_brute_force(oracle, dim, n):
    st, meta = problems.random_state(dim, n, seed=21, jitter=0.2)
    N = st["pos"].shape[0]
    b = oracle.box_boundaries(dim)
    pos, H, control, sub_of, _ = oracle.set_ghosts(dim, st["pos"], st["H"], b)
    off, nbr = oracle.connectivity(dim, N, pos, H)
    m, old2new = oracle.cull_map(N, pos.shape[0], off, nbr, control)
    keep = old2new >= 0
    assert np.all(keep[:N])
    # brute force: a ghost is seen by an internal node iff the pair criterion holds
    ns = dim * (dim + 1) // 2
    iu = np.triu_indices(dim)

    def full(h):
        M = np.zeros((dim, dim))
        M[iu] = h
        return M + np.triu(M, 1).T

    seen = np.zeros(pos.shape[0], dtype=bool)
    Hi = np.array([full(h) for h in H])
    for g in range(N, pos.shape[0]):
        r = pos[:N] - pos[g]
        ei = np.einsum("nab,nb->na", Hi[:N], r)
        eg = r @ Hi[g].T
        seen[g] = np.any(((ei ** 2).sum(1) <= 4.0) | ((eg ** 2).sum(1) <= 4.0))
    is_control = np.zeros(pos.shape[0], dtype=bool)
    is_control[control[control >= N]] = True
    assert np.array_equal(keep[N:], (seen | is_control)[N:])
    # kept nodes keep their order
    kept = np.nonzero(keep)[0]
    assert np.array_equal(old2new[kept], np.arange(m))
    assert ns > 0
