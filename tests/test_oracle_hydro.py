"""Pins the oracle's computeSPHSumMassDensity and evaluateDerivatives.

  * two-node head-on collision: every pair term of SPHHydroBase.cc:676-824 and
    MonaghanGingoldViscosity.cc:117-132 written out by hand (a dropped term, a
    wrong sign or an i<->j swap fails it);
  * linear velocity field v = A.x: the M-corrected velocity gradient equals A on
    every node with an invertible M (src/SPH/tests/testLinearVelocityGradient.py);
  * uniform state on a perfect lattice: interior DvDt, DepsDt, DvDx vanish;
  * conservation: sum_i m_i DvDt_i = 0 and sum_i m_i (v_i.DvDt_i + DepsDt_i) = 0
    when every node is internal (each pair term is antisymmetric);
  * sum density: interior lattice value equals an independent lattice sum;
  * the gather form (node-by-node, own frame) is bit-identical to the scatter
    form (reference loop) on random states, with tensile correction and Q
    multipliers - the identity the device kernels rely on;
  * GammaLawGas closed form.
"""
import math

import numpy as np
import pytest

from sphgpu import problems


def run(oracle, dim, state, n_int, gather=False, **kw):
    T = oracle.Table(dim, 1000)
    off, nbr = oracle.connectivity(dim, n_int, state["pos"], state["H"], 2.0)
    p = oracle.make_params(**kw)
    return T, off, nbr, oracle.eval_derivs(T, p, n_int, state, off, nbr, gather=gather)


def test_two_node_headon(oracle):
    dim = 3
    h, d, u, m, rho, eps, gamma = 0.1, 0.13, 0.7, 0.25, 1.3, 0.8, 5.0 / 3.0
    pos = np.array([[0.0, 0, 0], [d, 0, 0]])
    vel = np.array([[u, 0, 0], [-u, 0, 0]])
    H = problems.sym_identity(3, 2, 1.0 / h)
    P, cs = problems.gamma_law(np.full(2, rho), np.full(2, eps), gamma)
    st = dict(pos=pos, vel=vel, mass=np.full(2, m), rho=np.full(2, rho), eps=np.full(2, eps), P=P, cs=cs, H=H,
              omega=np.ones(2))
    T, off, nbr, out = run(oracle, dim, st, 2, XSPH=1)
    assert list(nbr) == [1, 0]
    Hdet = h ** -3
    eta = d / h
    W, gW = T.W(eta, Hdet), T.gradW(eta, Hdet)
    W0 = T.W(0.0, 1.0)
    gradWi = np.array([-gW / h, 0, 0])          # i = node 0, r_ij = -d x_hat
    vij = np.array([2 * u, 0, 0])
    mu = (2 * u * (-d / h)) / (eta ** 2 + 1e-2)
    c = cs[0]
    e = -1.0 * c * min(0.0, mu) + 1.0 * min(0.0, mu) ** 2
    QP = e / rho
    Qacc = 0.5 * QP * gradWi
    Prho = P[0] / rho ** 2
    delta = Prho * gradWi * 2 + 2 * Qacc
    workQ = vij @ Qacc
    r = 1e-13
    assert out["DvDt"][0] == pytest.approx(-m * delta, rel=r)
    assert out["DvDt"][1] == pytest.approx(m * delta, rel=r)
    assert out["DepsDt"][0] == pytest.approx(m * (Prho * vij @ gradWi + workQ), rel=r)
    assert out["DepsDt"][1] == pytest.approx(m * (Prho * vij @ gradWi + workQ), rel=r)
    assert out["maxViscousPressure"][0] == pytest.approx(rho * rho * abs(QP), rel=r)
    assert out["effViscousPressure"][0] == pytest.approx(m / rho * rho * rho * abs(QP) * W, rel=r)
    assert out["viscousWork"][0] == pytest.approx(m * workQ, rel=r)
    assert out["rhoSum"][0] == pytest.approx(m * W + m * W0 * Hdet, rel=r)
    assert out["normalization"][0] == pytest.approx(m / rho * W + m / rho * W0 * Hdet, rel=r)
    # one neighbour <= 2^nu: no M correction, DvDx /= rho
    DvDx = -(m * np.outer(vij, gradWi)) / rho
    assert out["DvDx"][0].reshape(3, 3) == pytest.approx(DvDx, rel=r, abs=1e-300)
    assert out["DrhoDt"][0] == pytest.approx(-rho * np.trace(DvDx), rel=r)
    M = -m * np.outer(-d * np.array([1.0, 0, 0]), gradWi)
    assert out["M"][0].reshape(3, 3) == pytest.approx(M, rel=r, abs=1e-300)
    w = 0.5 * (m / rho * W + m / rho * W)
    wsum = w + Hdet * m / rho * W0
    assert out["XSPHWeightSum"][0] == pytest.approx(wsum, rel=r)
    assert out["XSPHDeltaV"][0] == pytest.approx(-w * vij, rel=r)
    assert out["DxDt"][0] == pytest.approx(vel[0] - w * vij / wsum, rel=r)
    assert out["weightedNeighborSum"][0] == pytest.approx((abs(gW) / Hdet) ** (1 / 3), rel=r)
    thpt = np.outer([-d, 0, 0], [-d, 0, 0]) / d ** 6 * (gradWi @ gradWi) / Hdet ** 2
    m2 = out["massSecondMoment"][0]
    assert m2[0] == pytest.approx(thpt[0, 0], rel=r) and np.all(m2[1:] == 0)
    tr = np.trace(DvDx)
    assert out["DHDt"][0] == pytest.approx([-(1 / h) / 3 * tr, 0, 0, -(1 / h) / 3 * tr, 0, -(1 / h) / 3 * tr],
                                           rel=r, abs=1e-300)
    assert out["ncalc"] == 1


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_linear_velocity_gradient(oracle, dim):
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (dim, dim))
    n = {1: 64, 2: 20, 3: 9}[dim]
    st, meta = problems.linear_velocity(dim, n, A)
    T, off, nbr, out = run(oracle, dim, st, meta["n_internal"])
    DvDx = out["DvDx"].reshape(-1, dim, dim)
    # the correction applies where numNeighbors > 2^nu (SPHHydroBase.cc:847-851)
    corrected = np.diff(off) > 2 ** dim
    assert corrected.sum() > 0.9 * corrected.size
    err = np.abs(DvDx[corrected] - A[None]).max()
    assert err < 1e-11, err
    assert not np.allclose(DvDx[~corrected], A[None]) or (~corrected).sum() == 0


def test_uniform_lattice_interior(oracle):
    dim, n = 3, 12
    pos, mass, H, dx = problems.lattice(dim, n, jitter=0.0)
    N = pos.shape[0]
    rho, eps = np.ones(N), np.full(N, 2.0)
    P, cs = problems.gamma_law(rho, eps, 5 / 3)
    st = dict(pos=pos, vel=np.zeros((N, 3)), mass=mass, rho=rho, eps=eps, P=P, cs=cs, H=H, omega=np.ones(N))
    T, off, nbr, out = run(oracle, dim, st, N)
    interior = np.all((pos > 3.5 * dx) & (pos < 1 - 3.5 * dx), axis=1)
    scale = np.abs(out["DvDt"]).max()       # boundary nodes feel the free surface
    assert scale > 0
    assert np.abs(out["DvDt"][interior]).max() < 1e-13 * scale
    assert np.all(out["DepsDt"][interior] == 0)
    assert np.all(out["DvDx"][interior] == 0)
    # interior rhoSum: independent lattice sum with the same table
    Hdet = (1 / (1.51 * dx)) ** 3
    r = np.arange(-4, 5)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3) * dx
    eta = np.sqrt((g ** 2).sum(1)) / (1.51 * dx)
    ref = sum(mass[0] * T.W(e, Hdet) for e in eta)
    assert np.allclose(out["rhoSum"][interior], ref, rtol=1e-13, atol=0)
    # weighted neighbour sum: (sum_j |W'(eta_j)| Hdet / Hdet)^(1/3) over the lattice shell
    nz = eta > 0
    gw = np.array([abs(T.gradW(e, Hdet)) for e in eta[nz]])
    wns = (gw.sum() / Hdet) ** (1 / 3)
    assert np.allclose(out["weightedNeighborSum"][interior], wns, rtol=1e-13, atol=0)
    # second moment: sum_j (r r)/r^6 |gradW_j|^2 / Hdet^2 (diagonal, equal entries by symmetry)
    r2 = (g[nz] ** 2).sum(1)
    m2 = ((g[nz, 0] ** 2) / r2 ** 3 * (gw / (1.51 * dx)) ** 2).sum() / Hdet ** 2
    m2nd = out["massSecondMoment"].reshape(N, 6)[interior]
    assert np.allclose(m2nd[:, [0, 3, 5]], m2, rtol=1e-12, atol=0)
    assert np.abs(m2nd[:, [1, 2, 4]]).max() < 1e-12 * m2
    # on the lattice the table's own nPerh is recovered, so the ideal H is H
    # (to the nperh table's interpolation) and DHDt vanishes with tr DvDx
    hid = out["Hideal"].reshape(N, 6)[interior]
    assert np.allclose(hid[:, [0, 3, 5]], 1 / (1.51 * dx), rtol=2e-3, atol=0)
    assert np.all(out["DHDt"].reshape(N, 6)[interior] == 0)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_conservation(oracle, dim):
    n = {1: 80, 2: 16, 3: 7}[dim]
    st, meta = problems.random_state(dim, n, seed=11, h_var=0.1)
    T, off, nbr, out = run(oracle, dim, st, meta["n_internal"])
    m = st["mass"]
    mom = (m[:, None] * out["DvDt"].reshape(-1, dim)).sum(0)
    momscale = (np.abs(m[:, None] * out["DvDt"].reshape(-1, dim))).sum(0)
    assert np.all(np.abs(mom) <= 1e-13 * momscale)
    e = m * ((st["vel"] * out["DvDt"].reshape(-1, dim)).sum(1) + out["DepsDt"])
    escale = np.abs(m * (st["vel"] * out["DvDt"].reshape(-1, dim)).sum(1)).sum() + np.abs(m * out["DepsDt"]).sum()
    assert abs(e.sum()) <= 1e-13 * escale


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("variant", ["plain", "tensile", "multipliers", "quadratic", "total_energy"])
def test_gather_matches_scatter_bitwise(oracle, dim, variant):
    n = {1: 60, 2: 14, 3: 6}[dim]
    st, meta = problems.random_state(dim, n, seed=5 + dim, h_var=0.15)
    N = meta["n_internal"]
    n_int = N - N // 6     # a tail of nodes without rows stands in for ghosts
    kw = {}
    rng = np.random.default_rng(9)
    if variant == "tensile":
        kw["epsTensile"] = 0.3
        st["P"] = st["P"] - 1.0   # some negative pressures
    if variant == "multipliers":
        st["fCl"] = rng.uniform(0.5, 1.5, N)
        st["fCq"] = rng.uniform(0.5, 1.5, N)
        st["fshear"] = rng.uniform(0.2, 1.0, N)
    if variant == "quadratic":
        kw.update(linearInExpansion=1, quadraticInExpansion=1)
    if variant == "total_energy":
        kw.update(evolveTotalEnergy=1, XSPH=0, correctVelocityGradient=0)
    T = oracle.Table(dim, 1000)
    off, nbr = oracle.connectivity(dim, n_int, st["pos"], st["H"], 2.0)
    p = oracle.make_params(**kw)
    sc = oracle.eval_derivs(T, p, n_int, st, off, nbr, gather=False)
    ga = oracle.eval_derivs(T, p, n_int, st, off, nbr, gather=True)
    for f in oracle.OUTPUT_FIELDS:
        a, b = sc[f], ga[f]
        assert np.array_equal(a, b), (f, np.abs(a - b).max())
    # sum density, gather vs scatter
    r1 = oracle.sum_density(T, n_int, st["pos"], st["mass"], st["H"], off, nbr, gather=False)
    r2 = oracle.sum_density(T, n_int, st["pos"], st["mass"], st["H"], off, nbr, gather=True)
    assert np.array_equal(r1, r2)


def test_gamma_law(oracle):
    rho = np.array([0.5, 1.0, 2.0])
    eps = np.array([1e-6, 1.0, 3.0])
    P, cs = oracle.gamma_law(rho, eps, 1.4)
    assert P == pytest.approx(0.4 * rho * eps, rel=1e-15)
    assert cs == pytest.approx(np.sqrt(1.4 * 0.4 * eps), rel=1e-15)
    P2, _ = oracle.gamma_law(np.array([1.0]), np.array([-1.0]), 1.4, Pmin=0.0)
    assert P2[0] == 0.0


def test_omega_gradh_lattice(oracle):
    """computeSPHOmegaGradhCorrection on a uniform lattice interior equals the
    independent lattice sum -sum(eta W')/(nu (W(0) + sum W)); a node without
    neighbours gets 1."""
    dim, n = 3, 12
    pos, mass, H, dx = problems.lattice(dim, n)
    N = pos.shape[0]
    off, nbr = oracle.connectivity(dim, N, pos, H)
    T = oracle.Table(dim, 1000)
    om = oracle.omega_gradh(T, N, pos, H, off, nbr)
    interior = np.all((pos > 3.5 * dx) & (pos < 1 - 3.5 * dx), axis=1)
    Hdet = (1 / (1.51 * dx)) ** 3
    r = np.arange(-4, 5)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3) * dx
    eta = np.sqrt((g ** 2).sum(1)) / (1.51 * dx)
    nz = (eta > 0) & (eta < 2.0)
    Wsum = T.W(0.0, Hdet) + sum(T.W(e, Hdet) for e in eta[nz])
    gsum = sum(e * T.gradW(e, Hdet) for e in eta[nz])
    assert np.allclose(om[interior], -gsum / (3 * Wsum), rtol=1e-12, atol=0)
    assert np.all(om > 0)
    # an isolated node
    p1 = np.array([[0.0, 0.0, 0.0], [5.0, 5.0, 5.0]])
    H1 = problems.sym_identity(3, 2, 1.0)
    o1, n1 = oracle.connectivity(3, 2, p1, H1)
    assert np.array_equal(oracle.omega_gradh(T, 2, p1, H1, o1, n1), [1.0, 1.0])


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_pair_accelerations(oracle, dim):
    """pairAccelerations (SPHHydroBase.cc:794-797): each node's entries sum to
    its DvDt in the order they were pushed (the same bits), and the two ends of
    a pair carry equal and opposite momentum, m_i pa_i = -m_j pa_j."""
    n = {1: 60, 2: 14, 3: 6}[dim]
    st, meta = problems.random_state(dim, n, seed=19, h_var=0.1, vel_scale=0.3)
    N = meta["n_internal"]
    T = oracle.Table(dim, 1000)
    off, nbr = oracle.connectivity(dim, N, st["pos"], st["H"], 2.0)
    out = oracle.eval_derivs(T, oracle.make_params(), N, st, off, nbr, pair_acc=True)
    pa, m = out["pairAcc"], st["mass"]
    dv = out["DvDt"].reshape(N, dim)
    for i in range(N):
        acc = np.zeros(dim)
        for e in range(off[i], off[i + 1]):
            acc = acc + pa[e]
        assert np.array_equal(acc, dv[i]), i
    where = {(i, int(nbr[e])): e for i in range(N) for e in range(off[i], off[i + 1])}
    checked = 0
    for (i, j), e in where.items():
        if j > i and (j, i) in where:
            a, b = m[i] * pa[e], -m[j] * pa[where[(j, i)]]
            assert np.allclose(a, b, rtol=1e-14, atol=1e-14 * np.abs(a).max()), (i, j)
            checked += 1
    assert checked > 0


def test_compatible_depsdt_split(oracle):
    """SpecificThermalEnergyPolicy on one pair: the work du = (v_j12 - v_i12).pa_i
    is shared by w_i = |DepsDt0_i|/(|DepsDt0_i| + |DepsDt0_j|), the owner's
    share specific to its mass and the rest to the other node's; the total
    m_i de_i + m_j de_j is the pair's work m_i du whatever the split."""
    dim = 3
    off = np.array([0, 1, 2], dtype=np.int64)
    nbr = np.array([1, 0], dtype=np.int32)
    mass = np.array([2.0, 0.5])
    vel = np.array([[0.3, -0.1, 0.2], [-0.4, 0.25, 0.0]])
    acc = np.array([[1.0, 0.5, -2.0], [-4.0, -2.0, 8.0]])
    pa = np.array([[1.0, 0.5, -2.0], [-4.0, -2.0, 8.0]])      # m0 pa0 = -m1 pa1
    mult = 0.02
    v12 = vel + acc * (0.5 * mult)
    du = ((v12[1] - v12[0]) * pa[0]).sum()
    for d0, d1, w0 in [(3.0, 1.0, 0.75), (1.0, 0.0, 1.0), (0.0, 2.0, 0.0), (-2.0, 2.0, 0.5)]:
        de = oracle.compatible_depsdt(dim, 2, off, nbr, mass, vel, acc, np.array([d0, d1]), pa, mult)
        assert de[0] == pytest.approx(w0 * du, rel=1e-14, abs=1e-15 * abs(du))
        assert de[1] == pytest.approx((1.0 - w0) * du * mass[0] / mass[1], rel=1e-14, abs=1e-15 * abs(du))
        assert mass[0] * de[0] + mass[1] * de[1] == pytest.approx(mass[0] * du, rel=1e-14)
