"""Oracle checks for the time step (SynchronousRK2::step, State::update
policies, GenericHydro::dt, boundary enforcement).

Pinned by closed forms: the CFL vote of a lattice at rest is cfl*dx/c; the
step grows by dtGrowth from the integrator's initial last step (1e-5); a
uniform lattice with periodic or reflecting ghosts stays at rest; momentum is
conserved in a periodic box; energy drifts only at the truncation level; a
node behind a plane is mirrored or wrapped by the exact maps."""
import numpy as np
import pytest

from sphgpu import problems


@pytest.fixture(scope="module")
def integ():
    import integrator
    return integrator


def test_dt_vote_at_rest(oracle, integ):
    dim, n = 3, 6
    pos, mass, H, dx = problems.lattice(dim, n)
    N = pos.shape[0]
    st = dict(pos=pos, vel=np.zeros((N, dim)), mass=mass, rho=np.ones(N), eps=np.full(N, 2.0), H=H)
    r = integ.RK2(dim, st, [], 5.0 / 3.0)
    off, nbr = oracle.connectivity(dim, N, pos, H)
    full = dict(r.st)
    vote = integ.hydro_dt(dim, N, full, r.prev, off, nbr, 1.51, 0.25)
    nodeScale = 1.0 / (1.0 / (1.51 * dx)) / 1.51
    assert vote == pytest.approx(0.25 * nodeScale / full["cs"][0], rel=1e-15)
    assert nodeScale == pytest.approx(dx, rel=1e-15)


@pytest.mark.parametrize("kind", ["periodic", "reflecting"])
def test_lattice_at_rest_stays_at_rest(oracle, integ, kind):
    dim, n = 3, 6
    pos, mass, H, dx = problems.lattice(dim, n)
    N = pos.shape[0]
    st = dict(pos=pos, vel=np.zeros((N, dim)), mass=mass, rho=np.ones(N), eps=np.ones(N), H=H)
    r = integ.RK2(dim, st, oracle.box_boundaries(dim, kind=kind), 1.4)
    dts = [r.step() for _ in range(3)]
    assert dts == [2e-5, 4e-5, 8e-5]                 # dtGrowth * lastDt, from lastDt = 1e-5
    assert np.abs(r.st["vel"]).max() < 1e-15
    assert np.abs(r.st["pos"] - pos).max() < 1e-15
    assert np.allclose(r.st["rho"], r.st["rho"][0], rtol=1e-13, atol=0)


def test_periodic_momentum_and_energy(oracle, integ):
    st, meta = problems.periodic_sinusoid(6)
    r = integ.RK2(3, st, oracle.box_boundaries(3, kind="periodic"), 1.4)
    m = st["mass"]
    p0 = (m[:, None] * st["vel"]).sum(0)
    E0 = (0.5 * m * (st["vel"] ** 2).sum(1) + m * st["eps"]).sum()
    for _ in range(4):
        r.step()
    v = r.st["vel"]
    p1 = (m[:, None] * v).sum(0)
    scale = (m[:, None] * np.abs(v)).sum()
    assert np.all(np.abs(p1 - p0) < 1e-13 * scale)
    E1 = (0.5 * m * (v ** 2).sum(1) + m * r.st["eps"]).sum()
    assert abs(E1 - E0) < 1e-7 * E0


def test_enforce_boundaries_maps(oracle):
    b = oracle.box_boundaries(3)
    pos = np.array([[-0.01, 0.5, 0.5], [0.5, 1.02, 0.5], [0.3, 0.4, 0.5]])
    vel = np.array([[-1.0, 0.2, 0.3], [0.1, 2.0, 0.3], [0.1, 0.2, 0.3]])
    H = problems.sym_identity(3, 3, 40.0)
    p, v = pos.copy(), vel.copy()
    oracle.enforce_boundaries(3, 3, b, p, v, H)
    assert np.allclose(p[0], [0.01, 0.5, 0.5], atol=1e-16) and np.array_equal(v[0], [1.0, 0.2, 0.3])
    assert np.allclose(p[1], [0.5, 0.98, 0.5], atol=1e-15) and np.array_equal(v[1], [0.1, -2.0, 0.3])
    assert np.array_equal(p[2], pos[2]) and np.array_equal(v[2], vel[2])
    bp = oracle.box_boundaries(3, kind="periodic")
    p, v = pos.copy(), vel.copy()
    oracle.enforce_boundaries(3, 3, bp, p, v, H)
    assert np.allclose(p[0], [0.99, 0.5, 0.5], atol=1e-15) and np.array_equal(v[0], vel[0])
    assert np.allclose(p[1], [0.5, 0.02, 0.5], atol=1e-15)


def test_converging_flow_steps(oracle, integ):
    """Cold converging flow in a reflecting box: steps grow by dtGrowth until
    the CFL vote binds; energy is kept to the truncation level."""
    st, meta = problems.converging_flow(6)
    r = integ.RK2(3, st, oracle.box_boundaries(3), 5.0 / 3.0)
    m = st["mass"]
    E0 = (0.5 * m * (st["vel"] ** 2).sum(1) + m * st["eps"]).sum()
    dts = [r.step() for _ in range(5)]
    assert dts[:3] == [2e-5, 4e-5, 8e-5]
    E1 = (0.5 * m * (r.st["vel"] ** 2).sum(1) + m * r.st["eps"]).sum()
    assert abs(E1 - E0) < 1e-6 * E0
    assert r.t == pytest.approx(sum(dts), rel=1e-15)


@pytest.mark.parametrize("kind", ["periodic", "none"])
def test_compatible_energy_conserved(oracle, integ, kind):
    """compatibleEnergyEvolution: the thermal energy takes exactly the work of
    the pair forces, so sum m (v^2/2 + eps) holds to round-off over the steps
    where the plain update drifts at the truncation level."""
    if kind == "periodic":
        st, _ = problems.periodic_sinusoid(6)
        b = oracle.box_boundaries(3, kind="periodic")
    else:
        st, _ = problems.random_state(3, 6, seed=23, jitter=0.1, vel_scale=0.5, h_var=0.0)
        b = []
    m = st["mass"]
    E0 = (0.5 * m * (st["vel"] ** 2).sum(1) + m * st["eps"]).sum()
    drift = {}
    for compatible in (False, True):
        r = integ.RK2(3, dict(st), b, 1.4, compatible=compatible)
        for _ in range(4):
            r.step()
        v = r.st["vel"]
        E1 = (0.5 * m * (v ** 2).sum(1) + m * r.st["eps"]).sum()
        drift[compatible] = abs(E1 - E0) / E0
    assert drift[True] <= 1e-14
    assert drift[False] > 10 * drift[True]


def test_iterate_ideal_h_fixed_point(oracle, integ):
    """iterateIdealH on a periodic lattice: from a smoothing scale 30% too
    large and from one 20% too small the iteration reaches the same H, equal on
    every node, and there the lattice's equivalent nodes per smoothing scale
    (TableKernel::equivalentNodesPerSmoothingScale of the root-nu sum of
    |W'|) is nPerh, the fixed point of SPHSmoothingScale::idealSmoothingScale."""
    dim, n = 2, 20
    pos, mass, H, dx = problems.lattice(dim, n)
    N = pos.shape[0]
    b = oracle.box_boundaries(dim, kind="periodic")
    res = []
    for f in (1.3, 0.8):
        st = dict(pos=pos, mass=mass, rho=np.ones(N), H=H * f)
        H1, it, md = integ.iterate_ideal_h(dim, st, b)
        assert md <= 1.0e-10 and it < 100
        assert np.abs(H1[:, 0] - H1[0, 0]).max() <= 1e-13 * H1[0, 0]
        res.append(H1[0, 0])
    assert res[0] == pytest.approx(res[1], rel=1e-9)
    # fixed point: sum |W'(|H x_ij|)| over the lattice neighbours of one node
    T = oracle.Table(dim, 1000)
    h = res[0]
    d = (pos - pos[0] + 0.5) % 1.0 - 0.5
    eta = h * np.sqrt((d ** 2).sum(1))
    eta = eta[(eta > 0) & (eta < 2.0)]
    w = sum(abs(T.L.orc_table_grad(T.h, float(e), 1.0)) for e in eta)
    nperh = T.L.orc_table_nperh_from_wsum(T.h, float(np.sqrt(w)))
    assert nperh == pytest.approx(1.51, rel=1e-8)


def test_iterate_ideal_h_jittered_box(oracle, integ):
    """A jittered lattice in a reflecting box with a spread of smoothing
    scales converges below the tolerance, every change computed with the
    ghosts of the current H."""
    st, meta = problems.random_state(2, 14, seed=77, jitter=0.2, h_var=0.3)
    H1, it, md = integ.iterate_ideal_h(2, {k: st[k] for k in ("pos", "mass", "rho", "H")},
                                       oracle.box_boundaries(2), max_iterations=100, tolerance=1e-10)
    assert md <= 1e-10 and 2 < it < 100
    assert np.all(H1[:, 1] == 0.0) and np.all(H1[:, 0] == H1[:, 2])
