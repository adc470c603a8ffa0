"""The oracle with several NodeLists, pinned by what the single-NodeList
oracle (itself pinned) and independent sums fix.

The reference forms pairs across NodeLists exactly as within one
(ConnectivityMap over every NodeList, pair ownership by
ConnectivityMapInline.hh:169-186), and differs only in
 - evaluateDerivatives: rhoSum and normalization take same-NodeList pairs
   only (SPHHydroBase.cc:746-751); every other term is the same;
 - computeSPHSumMassDensity: a pair across NodeLists counts only with
   sumOverAllNodeLists, and then each end takes its own mass times the other
   end's kernel (computeSPHSumMassDensity.cc:77-100);
 - ghosts: each NodeList selects its own controls with its own hmax
   (PlanarBoundary.cc:112-182).
"""
import numpy as np
import pytest

from sphgpu import problems


def _two_lists(dim, n, seed, mass_ratio=1.0, boundaries=True, kind="reflecting"):
    st, _ = problems.random_state(dim, n, seed=seed, jitter=0.15, h_var=0.15, vel_scale=0.3)
    st = {k: np.array(v, dtype=np.float64) for k, v in st.items()}
    N = st["pos"].shape[0]
    # NodeList 1: the nodes right of the middle, in a contiguous block after NodeList 0
    right = st["pos"][:, 0] > 0.5
    order = np.concatenate([np.flatnonzero(~right), np.flatnonzero(right)])
    st = {k: v[order] for k, v in st.items()}
    lists = np.concatenate([np.zeros((~right).sum(), dtype=np.int32), np.ones(right.sum(), dtype=np.int32)])
    st["mass"] = st["mass"] * np.where(lists == 1, mass_ratio, 1.0)
    return st, lists, N


def _pipeline(oracle, dim, st, lists, kind):
    b = oracle.box_boundaries(dim, kind=kind)
    full, off, nbr, info = oracle.ghost_pipeline(dim, st, b, 2.0, True, lists=lists)
    return full, off, nbr, info


@pytest.mark.parametrize("dim,n", [(1, 60), (2, 12), (3, 6)])
def test_scatter_equals_gather_with_nodelists(oracle, dim, n):
    """The gather form in the scatter's delivery order reproduces the
    reference-order scatter bit for bit with two NodeLists and their ghosts."""
    st, lists, N = _two_lists(dim, n, seed=100 + dim, mass_ratio=2.5)
    P, c = problems.gamma_law(st["rho"], st["eps"], 1.4)
    st.update(P=P, cs=c, omega=np.ones(N))
    full, off, nbr, info = _pipeline(oracle, dim, st, lists, "reflecting")
    T = oracle.Table(dim, 1000)
    for sum_all in (True, False):
        a = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr, lists=full["list"],
                               sum_all=sum_all)
        b = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr, gather=True,
                               lists=full["list"], sum_all=sum_all)
        assert np.array_equal(a, b), sum_all
    p = oracle.make_params()
    a = oracle.eval_derivs(T, p, N, full, off, nbr, lists=full["list"])
    b = oracle.eval_derivs(T, p, N, full, off, nbr, gather=True, lists=full["list"])
    for f in oracle.OUTPUT_FIELDS:
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("dim,n", [(2, 12), (3, 6)])
def test_one_nodelist_label_is_the_single_list_oracle(oracle, dim, n):
    st, _, N = _two_lists(dim, n, seed=7)
    P, c = problems.gamma_law(st["rho"], st["eps"], 1.4)
    st.update(P=P, cs=c, omega=np.ones(N))
    zeros = np.zeros(N, dtype=np.int32)
    full, off, nbr, _ = _pipeline(oracle, dim, st, zeros, "periodic")
    full1, off1, nbr1, _ = oracle.ghost_pipeline(dim, st, oracle.box_boundaries(dim, kind="periodic"), 2.0, True)
    assert np.array_equal(full["pos"], full1["pos"]) and np.array_equal(nbr, nbr1)
    T = oracle.Table(dim, 1000)
    p = oracle.make_params()
    a = oracle.eval_derivs(T, p, N, full, off, nbr, lists=full["list"])
    b = oracle.eval_derivs(T, p, N, full1, off1, nbr1)
    for f in oracle.OUTPUT_FIELDS:
        assert np.array_equal(a[f], b[f]), f


def _cross_sums(T, dim, N, full, off, nbr, lists):
    """Independent numpy sums over the pairs across NodeLists: for each
    internal node x, sum over cross neighbours y of W(|H_x r|, det H_x) m_y
    (the rhoSum terms), of m_x/rho_x W(|H_x r|) (normalization), of
    m_y W(|H_y r|, det H_y) and of m_x W(|H_y r|, det H_y) (density terms)."""
    ns = dim * (dim + 1) // 2
    out = np.zeros((N, 4))
    iso = np.allclose(full["H"][:, 1:] if dim > 1 else 0.0, 0.0) if dim == 2 else None
    for x in range(N):
        for y in nbr[off[x]:off[x + 1]]:
            if lists[y] == lists[x]:
                continue
            r = full["pos"][x] - full["pos"][y]

            def W(h):
                Hm = np.zeros((dim, dim))
                k = 0
                for a in range(dim):
                    for b in range(a, dim):
                        Hm[a, b] = Hm[b, a] = h[k]
                        k += 1
                return T.W(np.linalg.norm(Hm @ r), np.linalg.det(Hm))
            Wx, Wy = W(full["H"][x]), W(full["H"][y])
            out[x] += [full["mass"][y] * Wx, full["mass"][x] / full["rho"][x] * Wx, full["mass"][y] * Wy,
                       full["mass"][x] * Wy]
    return out


@pytest.mark.parametrize("dim,n", [(1, 50), (2, 10)])
def test_nodelist_rules_against_independent_cross_sums(oracle, dim, n):
    """Two NodeLists of different masses: every derivative field but rhoSum and
    normalization equals the one-NodeList result (summation order only);
    those two lack exactly the cross-NodeList terms; the density differs by
    the cross terms' mass rule (sumOverAllNodeLists) or lacks them."""
    st, lists, N = _two_lists(dim, n, seed=30 + dim, mass_ratio=3.0)
    P, c = problems.gamma_law(st["rho"], st["eps"], 1.4)
    st.update(P=P, cs=c, omega=np.ones(N))
    full, off, nbr, _ = _pipeline(oracle, dim, st, lists, "reflecting")
    one = {k: v for k, v in full.items() if k != "list"}
    T = oracle.Table(dim, 1000)
    p = oracle.make_params()
    multi = oracle.eval_derivs(T, p, N, full, off, nbr, lists=full["list"])
    single = oracle.eval_derivs(T, p, N, one, off, nbr)
    cross = _cross_sums(T, dim, N, full, off, nbr, full["list"])
    for f in oracle.OUTPUT_FIELDS:
        a, b = np.asarray(multi[f]), np.asarray(single[f])
        if f == "rhoSum":
            b = b - cross[:, 0]
        elif f == "normalization":
            b = b - cross[:, 1]
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-300), f
    rho1 = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr)
    rho_all = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr, lists=full["list"])
    rho_own = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr, lists=full["list"],
                                 sum_all=False)
    assert np.abs(rho_all - (rho1 - cross[:, 2] + cross[:, 3])).max() <= 1e-12 * rho1.max()
    assert np.abs(rho_own - (rho1 - cross[:, 2])).max() <= 1e-12 * rho1.max()
    assert np.abs(cross[:, 3] - cross[:, 2]).max() > 1e-3 * rho1.max()     # the mass rule matters here


def test_each_nodelist_makes_its_own_ghosts(oracle):
    """With per-NodeList hmax, NodeList L's ghosts are the ghosts its own
    nodes make alone, in the same order."""
    dim = 2
    st, lists, N = _two_lists(dim, 12, seed=9)
    st["H"][lists == 1] *= 0.7           # NodeList 1 has larger smoothing scales
    b = oracle.box_boundaries(dim)
    pos, H, control, sub_of, _, la = oracle.set_ghosts(dim, st["pos"], st["H"], b, 2.0, lists)
    for L in (0, 1):
        sel = np.flatnonzero(lists == L)
        posL, HL, _, _, _ = oracle.set_ghosts(dim, st["pos"][sel], st["H"][sel], b, 2.0)
        mine = la[N:] == L
        assert np.array_equal(pos[N:][mine], posL[sel.size:]), L
        assert np.array_equal(H[N:][mine], HL[sel.size:]), L


def test_integrator_one_nodelist_label_is_the_plain_integrator(oracle):
    """The oracle's time step with every node in NodeList 0 (and that
    NodeList's gamma) is the one-NodeList time step, bit for bit."""
    import integrator
    dim = 2
    st, _, N = _two_lists(dim, 10, seed=11)
    base = {k: st[k] for k in ("pos", "vel", "mass", "rho", "eps", "H")}
    b = oracle.box_boundaries(dim, kind="reflecting")
    a = integrator.RK2(dim, base, b, 1.4)
    c = integrator.RK2(dim, base, b, 2.0, lists=np.zeros(N, dtype=np.int32), gammas=(1.4,))
    for _ in range(3):
        assert a.step() == c.step()
    for k in ("pos", "vel", "eps", "rho", "H", "P", "cs"):
        assert np.array_equal(a.st[k], c.st[k]), k


@pytest.mark.parametrize("dim,n", [(1, 60), (2, 12), (3, 6)])
def test_omega_with_nodelists_has_no_nodelist_rule(oracle, dim, n):
    """The grad-h correction sums every pair alike across NodeLists
    (computeSPHOmegaGradhCorrection.cc:55-130): with two NodeLists it is the
    one-NodeList sum over the same rows, up to the order the owner loop walks
    them (ConnectivityMap order instead of ascending index), and one NodeList
    label changes nothing at all."""
    st, lists, N = _two_lists(dim, n, seed=120 + dim, mass_ratio=2.0)
    full, off, nbr, info = _pipeline(oracle, dim, st, lists, "reflecting")
    T = oracle.Table(dim, 1000)
    one = oracle.omega_gradh(T, N, full["pos"], full["H"], off, nbr)
    two = oracle.omega_gradh(T, N, full["pos"], full["H"], off, nbr, lists=full["list"])
    assert np.abs(two - one).max() <= 1e-14 * np.abs(one).max()
    same = oracle.omega_gradh(T, N, full["pos"], full["H"], off, nbr, lists=np.zeros_like(full["list"]))
    assert np.array_equal(same, one)


def test_integrator_gradh_with_nodelists_runs_the_correction(oracle):
    """The time step with grad-h and two NodeLists uses the correction: the
    states differ from the same step without it, and a single NodeList label
    gives the plain grad-h step bit for bit."""
    import integrator
    dim = 2
    st, lists, N = _two_lists(dim, 10, seed=13, mass_ratio=2.0)
    base = {k: st[k] for k in ("pos", "vel", "mass", "rho", "eps", "H")}
    b = oracle.box_boundaries(dim, kind="reflecting")
    g = integrator.RK2(dim, base, b, 1.4, gradh=True, lists=lists, gammas=(1.4, 1.4))
    p = integrator.RK2(dim, base, b, 1.4, gradh=False, lists=lists, gammas=(1.4, 1.4))
    for _ in range(2):
        g.step()
        p.step()
    assert not np.array_equal(g.st["vel"], p.st["vel"])
    a = integrator.RK2(dim, base, b, 1.4, gradh=True)
    c = integrator.RK2(dim, base, b, 1.4, gradh=True, lists=np.zeros(N, dtype=np.int32), gammas=(1.4,))
    for _ in range(2):
        assert a.step() == c.step()
    for k in ("pos", "vel", "eps", "rho", "H"):
        assert np.array_equal(a.st[k], c.st[k]), k


def test_compatible_energy_with_nodelists_conserved(oracle):
    """compatibleEnergyEvolution over two NodeLists (periodic box, gamma 1.4
    and 5/3): every pair's work goes to the thermal energies of its two ends,
    so sum m (v^2/2 + eps) holds to round-off; pairing an end with another
    entry's pair acceleration (a wrong row order) breaks it.  One NodeList
    label gives the plain compatible step bit for bit."""
    import integrator
    dim = 2
    st, lists, N = _two_lists(dim, 10, seed=17, mass_ratio=2.0)
    base = {k: st[k] for k in ("pos", "vel", "mass", "rho", "eps", "H")}
    b = oracle.box_boundaries(dim, kind="periodic")
    m = st["mass"]
    E0 = (0.5 * m * (st["vel"] ** 2).sum(1) + m * st["eps"]).sum()
    drift = {}
    for compatible in (False, True):
        r = integrator.RK2(dim, base, b, 1.4, compatible=compatible, lists=lists, gammas=(1.4, 5.0 / 3.0))
        for _ in range(4):
            r.step()
        E1 = (0.5 * m * (r.st["vel"] ** 2).sum(1) + m * r.st["eps"]).sum()
        drift[compatible] = abs(E1 - E0) / E0
    assert drift[True] <= 1e-14, drift
    assert drift[False] > 10 * drift[True], drift
    a = integrator.RK2(dim, base, b, 1.4, compatible=True)
    c = integrator.RK2(dim, base, b, 1.4, compatible=True, lists=np.zeros(N, dtype=np.int32), gammas=(1.4,))
    for _ in range(3):
        assert a.step() == c.step()
    for k in ("pos", "vel", "eps", "rho", "H"):
        assert np.array_equal(a.st[k], c.st[k]), k


@pytest.mark.parametrize("dim,n", [(1, 60), (2, 12)])
def test_ideal_h_weights_across_nodelists(oracle, dim, n):
    """iterateIdealH's zeroth moment weights a neighbour of another NodeList
    by m_j rho_i / (m_i rho_j) (iterateIdealH.cc:183-194): with one NodeList
    label the pass is the one-NodeList pass bit for bit; with two NodeLists
    whose m/rho agree the weight is 1 and the pass is the one-NodeList one up
    to the order of the sum; doubling one NodeList's mass alone changes it."""
    import integrator
    st, lists, N = _two_lists(dim, n, seed=160 + dim, mass_ratio=1.0)
    b = oracle.box_boundaries(dim, kind="reflecting")
    base = {k: st[k] for k in ("pos", "mass", "rho", "H")}
    one, it1, _ = integrator.iterate_ideal_h(dim, base, b, max_iterations=3)
    lab, it2, _ = integrator.iterate_ideal_h(dim, base, b, max_iterations=3, lists=np.zeros(N, dtype=np.int32))
    assert it1 == it2 and np.array_equal(one, lab)
    # m and rho both doubled in NodeList 1: weight m_j rho_i / (m_i rho_j) = 1
    same = dict(base, mass=base["mass"] * np.where(lists == 1, 2.0, 1.0),
                rho=base["rho"] * np.where(lists == 1, 2.0, 1.0))
    two, _, _ = integrator.iterate_ideal_h(dim, same, b, max_iterations=3, lists=lists)
    assert np.abs(two - one).max() <= 1e-14 * np.abs(one).max()
    heavy = dict(base, mass=base["mass"] * np.where(lists == 1, 2.0, 1.0))
    diff, _, _ = integrator.iterate_ideal_h(dim, heavy, b, max_iterations=3, lists=lists)
    assert np.abs(diff - one).max() > 1e-6 * np.abs(one).max()
