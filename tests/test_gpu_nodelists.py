"""Several NodeLists in one device node set against the oracle.

Two NodeLists of different masses (and, in the time steps, different gamma)
in a reflecting or periodic box: each NodeList makes its own ghosts, the
neighbour rows come in the order the reference's scatter delivers each
node's terms, rhoSum and normalization take same-NodeList pairs only, and
the summed density takes a pair across NodeLists only with
sumMassDensityOverAllNodeLists, each end with its own mass.  Bars as the
one-NodeList parity tests: reference row order bitwise (3-D cube-root
fields 1e-14), search row order 1e-12 relative to each field's largest
magnitude."""
import numpy as np
import pytest

from sphgpu import problems

pytestmark = pytest.mark.gpu

POW_FIELDS = ("weightedNeighborSum", "Hideal")
STATE = {"pos": "position", "vel": "velocity", "mass": "mass", "rho": "mass density",
         "eps": "specific thermal energy", "H": "H", "P": "pressure", "cs": "sound speed"}
STEP_FIELDS = {"pos": "position", "vel": "velocity", "eps": "specific thermal energy", "rho": "mass density",
               "H": "H", "P": "pressure", "cs": "sound speed"}


def two_lists(dim, n, seed, mass_ratio, gamma=(1.4, 1.4)):
    """NodeList 0 the nodes left of x = 0.5, NodeList 1 the others (heavier
    by mass_ratio, with larger smoothing scales), in that order."""
    st, _ = problems.random_state(dim, n, seed=seed, jitter=0.15, h_var=0.15, vel_scale=0.3)
    st = {k: np.array(v, dtype=np.float64) for k, v in st.items()}
    right = st["pos"][:, 0] > 0.5
    order = np.concatenate([np.flatnonzero(~right), np.flatnonzero(right)])
    st = {k: v[order] for k, v in st.items()}
    counts = [int((~right).sum()), int(right.sum())]
    lists = np.repeat(np.arange(2, dtype=np.int32), counts)
    st["mass"] = st["mass"] * np.where(lists == 1, mass_ratio, 1.0)
    st["H"][lists == 1] *= 0.85
    P, cs = np.zeros(lists.size), np.zeros(lists.size)
    for L, g in enumerate(gamma):
        P[lists == L], cs[lists == L] = problems.gamma_law(st["rho"][lists == L], st["eps"][lists == L], g)
    st.update(P=P, cs=cs, omega=np.ones(lists.size))
    return st, lists, counts


def device(sg, dim, st, counts, kind, order, gamma=None, sum_all=True):
    N = st["pos"].shape[0]
    s = sg.Sim(dim)
    s.set_row_order(order)
    s.set_table_bspline(1000)
    s.set_nodes(N, N)
    s.upload_state(st)
    s.set_node_lists(counts, gamma, sum_all)
    s.set_boundaries(sg.box_boundaries(dim, kind=kind))
    return s


def close(a, b, order, dim, f, tol=1e-12):
    scale = max(np.abs(b).max(), 1e-300)
    if order == 1:
        return np.abs(a - b).max() <= tol * scale
    if dim == 3 and f in POW_FIELDS:
        return np.abs(a - b).max() <= 1e-14 * scale
    return np.array_equal(a, b)


CASES = [(1, 80, "reflecting"), (2, 14, "reflecting"), (2, 14, "periodic"), (3, 8, "reflecting"),
         (3, 8, "periodic")]


@pytest.mark.parametrize("sum_all", [True, False], ids=["sumall", "ownlist"])
@pytest.mark.parametrize("order", [0, 1], ids=["refo", "searcho"])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[2]}{c[0]}d" for c in CASES])
def test_pass_with_nodelists(sphgpu, oracle, case, order, sum_all):
    dim, n, kind = case
    print(f"[nodelists-pass] dim={dim} n={n} {kind} order={order} sum_all={sum_all}", flush=True)
    st, lists, counts = two_lists(dim, n, seed=70 + dim, mass_ratio=2.5)
    N = lists.size
    bo = oracle.box_boundaries(dim, kind=kind)
    full, off, nbr, info = oracle.ghost_pipeline(dim, st, bo, 2.0, True, lists=lists)
    s = device(sphgpu, dim, st, counts, kind, order, sum_all=sum_all)
    s.set_ghost_nodes(2.0, cull=True)
    n_int, n_tot = s.node_counts()
    assert n_int == N and n_tot == full["pos"].shape[0]
    for k, f in STATE.items():
        assert np.array_equal(s.download(f).reshape(full[k].shape), full[k]), k
    control, sub = s.ghost_controls()
    assert np.array_equal(control, info["control"]) and np.array_equal(sub, info["sub_of"])
    # the ConnectivityMap download: rows ascending in caller index, whatever
    # order the device walks them in (the scatter's delivery order with
    # reference row order: the bitwise sums below hold it to that)
    o2, n2 = s.neighbors()
    assert np.array_equal(o2, off) and np.array_equal(n2, nbr)
    T = oracle.Table(dim, 1000)
    rho = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr, lists=full["list"],
                             sum_all=sum_all)
    s.sum_density()
    rho_all = np.array(full["rho"])
    rho_all[:N] = rho
    rho_all = oracle.apply_ghosts(dim, N, bo, info["control"], info["sub_of"], rho_all, 0)
    assert close(s.download("mass density"), rho_all, order, dim, "rho")
    full["rho"] = rho_all
    ref = oracle.eval_derivs(T, oracle.make_params(), N, full, off, nbr, lists=full["list"])
    s.evaluate_derivatives(sphgpu.hydro_params())
    for f in sphgpu.DERIV_ALIASES:
        assert close(s.derivative(f).reshape(ref[f].shape), ref[f], order, dim, f), f


def test_nodelist_rules_change_the_result(sphgpu):
    """The NodeList rules act on the device: rhoSum and the summed density of
    the same nodes differ between one and two NodeLists, and between the
    two density rules."""
    dim = 2
    st, lists, counts = two_lists(dim, 14, seed=3, mass_ratio=3.0)
    N = lists.size
    out = {}
    for name, c, sa in (("one", [N], True), ("all", counts, True), ("own", counts, False)):
        s = device(sphgpu, dim, st, c, "reflecting", 0, sum_all=sa)
        s.set_ghost_nodes(2.0, cull=True)
        s.sum_density()
        rho = s.download("mass density", rows=N)
        s.evaluate_derivatives(sphgpu.hydro_params())
        out[name] = (rho, s.derivative("rhoSum").copy())
    assert not np.array_equal(out["one"][0], out["all"][0])
    assert not np.array_equal(out["all"][0], out["own"][0])
    assert not np.array_equal(out["one"][1], out["all"][1])
    assert np.array_equal(out["all"][1], out["own"][1])   # rhoSum does not depend on the density rule


STEP_CASES = [(1, 80, "reflecting", 5), (2, 14, "periodic", 4), (3, 7, "reflecting", 3)]


@pytest.mark.parametrize("sum_all", [True, False], ids=["sumall", "ownlist"])
@pytest.mark.parametrize("case", STEP_CASES, ids=[f"{c[2]}{c[0]}d" for c in STEP_CASES])
def test_rk2_with_nodelists(sphgpu, oracle, case, sum_all):
    """SynchronousRK2 over two NodeLists with their own gamma: the same dt
    sequence and state as the oracle's integrator."""
    import integrator
    dim, n, kind, steps = case
    print(f"[nodelists-rk2] dim={dim} n={n} {kind} sum_all={sum_all}", flush=True)
    gammas = (1.4, 5.0 / 3.0)
    st, lists, counts = two_lists(dim, n, seed=90 + dim, mass_ratio=2.0, gamma=gammas)
    N = lists.size
    base = {k: st[k] for k in ("pos", "vel", "mass", "rho", "eps", "H")}
    ref = integrator.RK2(dim, base, oracle.box_boundaries(dim, kind=kind), gammas[0], lists=lists,
                         gammas=gammas, sum_all=sum_all)
    s = device(sphgpu, dim, base, counts, kind, sphgpu.ROW_ORDER_REFERENCE, gamma=gammas, sum_all=sum_all)
    s.eos_gamma_law(gammas[0])
    hp = sphgpu.hydro_params()
    ip = sphgpu.integrator_params(gamma=gammas[0])
    for k in range(steps):
        dref = ref.step()
        ddev = s.rk2_step(hp, ip)
        if dim < 3:
            assert ddev == dref, (k, ddev, dref)
        else:
            assert ddev == pytest.approx(dref, rel=1e-12), (k, ddev, dref)
    for key, f in STEP_FIELDS.items():
        a = s.download(f, rows=N).reshape(ref.st[key].shape)
        b = ref.st[key]
        if dim < 3:
            assert np.array_equal(a, b), key
        else:
            assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-300), key


def test_nodelist_argument_errors(sphgpu):
    dim = 2
    st, lists, counts = two_lists(dim, 10, seed=5, mass_ratio=1.0)
    N = lists.size
    s = device(sphgpu, dim, st, counts, "reflecting", 0)
    with pytest.raises(sphgpu.SphGpuError):
        s.set_node_lists([N, 1])                 # sizes do not add up
    with pytest.raises(sphgpu.SphGpuError):
        s.set_node_lists([N - 1, 1], gamma=[1.4, 1.0])
    s.set_ghost_nodes(2.0, cull=True)
    with pytest.raises(sphgpu.SphGpuError):
        s.set_node_lists(counts)                 # ghosts exist: internal rows only


GRADH_CASES = [(1, 80, "reflecting"), (2, 14, "periodic"), (3, 8, "reflecting")]


@pytest.mark.parametrize("order", [0, 1], ids=["refo", "searcho"])
@pytest.mark.parametrize("case", GRADH_CASES, ids=[f"{c[2]}{c[0]}d" for c in GRADH_CASES])
def test_omega_gradh_with_nodelists(sphgpu, oracle, case, order):
    """computeSPHOmegaGradhCorrection over two NodeLists: every pair alike,
    node i's terms in the owner loop's delivery order.  Reference row order
    value for value (ghost rows from their controls), search row order 1e-12;
    the pass that follows uses it."""
    dim, n, kind = case
    print(f"[nodelists-omega] dim={dim} n={n} {kind} order={order}", flush=True)
    st, lists, counts = two_lists(dim, n, seed=130 + dim, mass_ratio=2.0)
    N = lists.size
    bo = oracle.box_boundaries(dim, kind=kind)
    full, off, nbr, info = oracle.ghost_pipeline(dim, st, bo, 2.0, True, lists=lists)
    T = oracle.Table(dim, 1000)
    om = oracle.omega_gradh(T, N, full["pos"], full["H"], off, nbr, lists=full["list"])
    om_all = np.array(full["omega"])
    om_all[:N] = om
    om_all = oracle.apply_ghosts(dim, N, bo, info["control"], info["sub_of"], om_all, 0)
    s = device(sphgpu, dim, st, counts, kind, order)
    s.set_ghost_nodes(2.0, cull=True)
    s.omega_gradh()
    assert close(s.download("grad h corrections"), om_all, order, dim, "omega")
    # the derivative pass with the correction
    s.sum_density()
    rho_all = np.array(full["rho"])
    rho_all[:N] = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr, lists=full["list"])
    full["rho"] = oracle.apply_ghosts(dim, N, bo, info["control"], info["sub_of"], rho_all, 0)
    full["omega"] = om_all
    ref = oracle.eval_derivs(T, oracle.make_params(), N, full, off, nbr, lists=full["list"])
    s.evaluate_derivatives(sphgpu.hydro_params())
    for f in sphgpu.DERIV_ALIASES:
        assert close(s.derivative(f).reshape(ref[f].shape), ref[f], order, dim, f), f


@pytest.mark.parametrize("case", STEP_CASES[:2], ids=[f"{c[2]}{c[0]}d" for c in STEP_CASES[:2]])
def test_rk2_gradh_with_nodelists(sphgpu, oracle, case):
    """SynchronousRK2 with the grad-h correction over two NodeLists (gamma
    1.4 and 5/3): the oracle's dt sequence and state, omega included."""
    import integrator
    dim, n, kind, steps = case
    print(f"[nodelists-rk2-gradh] dim={dim} n={n} {kind}", flush=True)
    gammas = (1.4, 5.0 / 3.0)
    st, lists, counts = two_lists(dim, n, seed=140 + dim, mass_ratio=2.0, gamma=gammas)
    N = lists.size
    base = {k: st[k] for k in ("pos", "vel", "mass", "rho", "eps", "H")}
    ref = integrator.RK2(dim, base, oracle.box_boundaries(dim, kind=kind), gammas[0], lists=lists,
                         gammas=gammas, gradh=True)
    s = device(sphgpu, dim, base, counts, kind, sphgpu.ROW_ORDER_REFERENCE, gamma=gammas)
    s.eos_gamma_law(gammas[0])
    hp = sphgpu.hydro_params()
    ip = sphgpu.integrator_params(gamma=gammas[0], gradhCorrection=1)
    for k in range(steps):
        assert s.rk2_step(hp, ip) == ref.step(), k
    fields = dict(STEP_FIELDS, omega="grad h corrections")
    for key, f in fields.items():
        a = s.download(f, rows=N).reshape(ref.st[key].shape)
        assert np.array_equal(a, ref.st[key]), key


@pytest.mark.parametrize("case", STEP_CASES[:2], ids=[f"{c[2]}{c[0]}d" for c in STEP_CASES[:2]])
def test_rk2_compatible_with_nodelists(sphgpu, oracle, case):
    """SynchronousRK2 with compatibleEnergyEvolution over two NodeLists (gamma
    1.4 and 5/3): each node's pair terms replayed in the order the owner loop
    delivers them; the oracle's dt sequence and state value for value, and in
    the periodic box total energy to round-off."""
    import integrator
    dim, n, kind, steps = case
    print(f"[nodelists-rk2-compatible] dim={dim} n={n} {kind}", flush=True)
    gammas = (1.4, 5.0 / 3.0)
    st, lists, counts = two_lists(dim, n, seed=150 + dim, mass_ratio=2.0, gamma=gammas)
    N = lists.size
    base = {k: st[k] for k in ("pos", "vel", "mass", "rho", "eps", "H")}
    ref = integrator.RK2(dim, base, oracle.box_boundaries(dim, kind=kind), gammas[0], lists=lists,
                         gammas=gammas, compatible=True)
    s = device(sphgpu, dim, base, counts, kind, sphgpu.ROW_ORDER_REFERENCE, gamma=gammas)
    s.eos_gamma_law(gammas[0])
    hp = sphgpu.hydro_params(compatibleEnergyEvolution=1)
    ip = sphgpu.integrator_params(gamma=gammas[0])
    for k in range(steps):
        assert s.rk2_step(hp, ip) == ref.step(), k
    for key, f in STEP_FIELDS.items():
        a = s.download(f, rows=N).reshape(ref.st[key].shape)
        assert np.array_equal(a, ref.st[key]), key
    if kind == "periodic":
        m = st["mass"]
        v = s.download("velocity", rows=N).reshape(N, dim)
        e = s.download("specific thermal energy", rows=N)
        E0 = (0.5 * m * (st["vel"] ** 2).sum(1) + m * st["eps"]).sum()
        E1 = (0.5 * m * (v ** 2).sum(1) + m * e).sum()
        assert abs(E1 - E0) <= 1e-13 * E0


@pytest.mark.parametrize("case", [(1, 80, "reflecting"), (2, 14, "periodic"), (3, 7, "reflecting")],
                         ids=["reflecting1d", "periodic2d", "reflecting3d"])
def test_iterate_ideal_h_with_nodelists(sphgpu, oracle, case):
    """iterateIdealH over two NodeLists (the second twice as heavy): each
    NodeList's own ghosts, every node's full connectivity in ConnectivityMap
    order and neighbours of the other NodeList weighted by m_j rho_i /
    (m_i rho_j).  Same iteration count; H value for value in 1-D and 2-D,
    1e-12 in 3-D (cube root); the pass after it runs on the rows it leaves."""
    import integrator
    dim, n, kind = case
    print(f"[nodelists-idealh] dim={dim} n={n} {kind}", flush=True)
    st, lists, counts = two_lists(dim, n, seed=170 + dim, mass_ratio=2.0)
    N = lists.size
    Hr, itr, mdr = integrator.iterate_ideal_h(dim, {k: st[k] for k in ("pos", "mass", "rho", "H")},
                                              oracle.box_boundaries(dim, kind=kind), lists=lists)
    s = device(sphgpu, dim, st, counts, kind, sphgpu.ROW_ORDER_REFERENCE)
    it, md = s.iterate_ideal_h(sphgpu.hydro_params(), 100, 1.0e-10)
    assert it == itr, (it, itr)
    Hd = s.download("H", rows=N).reshape(Hr.shape)
    if dim < 3:
        assert np.array_equal(Hd, Hr)
        assert md == mdr
    else:
        assert np.abs(Hd - Hr).max() <= 1e-12 * np.abs(Hr).max()
    s.sum_density()
    s.evaluate_derivatives(sphgpu.hydro_params())


def rel_err(a, b):
    a, b = np.asarray(a, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()
    return (float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)),
            float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)))


MIXED_CASES = [(1, 80, "reflecting"), (2, 14, "periodic"), (3, 8, "reflecting")]


@pytest.mark.parametrize("sum_all", [True, False], ids=["sumall", "ownlist"])
@pytest.mark.parametrize("case", MIXED_CASES, ids=[f"{c[2]}{c[0]}d" for c in MIXED_CASES])
def test_mixed_pass_with_nodelists(sphgpu, oracle, case, sum_all):
    """Mixed precision over two NodeLists: the density sum's cross-NodeList
    rule and the same-NodeList rhoSum and normalization in the FP32 pair
    kernels, each field within 1e-5 of the FP64 oracle (relative L-inf and
    L2)."""
    dim, n, kind = case
    print(f"[nodelists-mixed] dim={dim} n={n} {kind} sum_all={sum_all}", flush=True)
    st, lists, counts = two_lists(dim, n, seed=180 + dim, mass_ratio=2.5)
    N = lists.size
    bo = oracle.box_boundaries(dim, kind=kind)
    full, off, nbr, info = oracle.ghost_pipeline(dim, st, bo, 2.0, True, lists=lists)
    T = oracle.Table(dim, 1000)
    rho = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr, lists=full["list"],
                             sum_all=sum_all)
    rho_all = np.array(full["rho"])
    rho_all[:N] = rho
    full["rho"] = oracle.apply_ghosts(dim, N, bo, info["control"], info["sub_of"], rho_all, 0)
    ref = oracle.eval_derivs(T, oracle.make_params(), N, full, off, nbr, lists=full["list"])
    s = device(sphgpu, dim, st, counts, kind, sphgpu.ROW_ORDER_SEARCH, sum_all=sum_all)
    s.set_precision(sphgpu.PRECISION_MIXED)
    s.set_ghost_nodes(2.0, cull=True)
    s.sum_density()
    bad = [("rho",) + rel_err(s.download("mass density", rows=N), rho)]
    s.evaluate_derivatives(sphgpu.hydro_params(precision=sphgpu.PRECISION_MIXED))
    for f in sphgpu.DERIV_ALIASES:
        bad.append((f,) + rel_err(s.derivative(f).reshape(ref[f].shape), ref[f]))
    bad = [b for b in bad if max(b[1:]) > 1e-5]
    assert not bad, bad


@pytest.mark.parametrize("case", STEP_CASES[:2], ids=[f"{c[2]}{c[0]}d" for c in STEP_CASES[:2]])
def test_rk2_mixed_with_nodelists(sphgpu, oracle, case):
    """Mixed-precision SynchronousRK2 steps over two NodeLists (gamma 1.4 and
    5/3): dt and state within 1e-5 of the FP64 oracle's."""
    import integrator
    dim, n, kind, steps = case
    print(f"[nodelists-rk2-mixed] dim={dim} n={n} {kind}", flush=True)
    gammas = (1.4, 5.0 / 3.0)
    st, lists, counts = two_lists(dim, n, seed=190 + dim, mass_ratio=2.0, gamma=gammas)
    N = lists.size
    base = {k: st[k] for k in ("pos", "vel", "mass", "rho", "eps", "H")}
    ref = integrator.RK2(dim, base, oracle.box_boundaries(dim, kind=kind), gammas[0], lists=lists, gammas=gammas)
    s = device(sphgpu, dim, base, counts, kind, sphgpu.ROW_ORDER_SEARCH, gamma=gammas)
    s.set_precision(sphgpu.PRECISION_MIXED)
    s.eos_gamma_law(gammas[0])
    hp = sphgpu.hydro_params(precision=sphgpu.PRECISION_MIXED)
    ip = sphgpu.integrator_params(gamma=gammas[0])
    for k in range(steps):
        dref = ref.step()
        ddev = s.rk2_step(hp, ip)
        assert ddev == pytest.approx(dref, rel=1e-5), (k, ddev, dref)
    for key, f in STEP_FIELDS.items():
        err = rel_err(s.download(f, rows=N).reshape(ref.st[key].shape), ref.st[key])
        assert max(err) <= 1e-5, (key, err)
