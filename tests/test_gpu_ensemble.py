"""Ensembles and the request loop on the device.

The contract: a member's results do not depend on the other members.  Each
member is checked against the same problem run alone as a plain simulation
(sphgpu_rk2_step with max_time = its end time), value for value, after the
ensemble has added and removed other members around it.  The plain
simulation's own parity with the oracle is covered by the integrator tests.
"""
import numpy as np
import pytest

from sphgpu import problems

pytestmark = pytest.mark.gpu

FIELDS = ("position", "velocity", "mass density", "specific thermal energy", "H")


def plain_run(sg, dim, st, gamma, t_end, hp, ip_kw, row_order, precision=0):
    N = st["mass"].shape[0]
    s = sg.Sim(dim)
    s.set_row_order(row_order)
    s.set_table_bspline(1000)
    s.set_nodes(N, N)
    s.upload_state({k: st[k] for k in ("pos", "vel", "mass", "rho", "eps", "H")})
    s.eos_gamma_law(gamma)
    s.set_boundaries(sg.box_boundaries(dim, kind="reflecting"))
    s.set_precision(precision)
    ip = sg.integrator_params(gamma=gamma, **ip_kw)
    steps = 0
    while s.clock()[0] < t_end:
        s.rk2_step(hp, ip, max_time=t_end)
        steps += 1
    out = {f: s.download(f, rows=N) for f in FIELDS}
    s.close()
    return out, steps


def ensemble(sg, dim, row_order, precision=0):
    e = sg.Sim(dim)
    e.set_row_order(row_order)
    e.set_table_bspline(1000)
    e.set_precision(precision)
    e.set_boundaries(sg.box_boundaries(dim, kind="reflecting"))
    e.ensemble_begin()
    return e


def run_ensemble(sg, e, members, hp, ip_kw, late=(), failed=None):
    """Step until every member finished; members in `late` join after the
    first step.  Returns {index: (fields, cycle)}; members the steps evicted
    go to `failed` ({index: (reason, cycle)})."""
    ip = sg.integrator_params(**ip_kw)
    ids = {}
    for q, m in enumerate(members):
        if q not in late:
            ids[e.ensemble_add(m["state"], m["gamma"], m["t_end"], sg.integrator_params(gamma=m["gamma"]))] = q
    res = {}
    first = True
    while ids:
        done = e.ensemble_step(hp, ip)
        for f in e.ensemble_failed():
            q = ids.pop(f["id"])
            assert failed is not None, (q, f)
            failed[q] = (f["reason"], f["cycle"])
        if first:
            for q in late:
                m = members[q]
                ids[e.ensemble_add(m["state"], m["gamma"], m["t_end"], sg.integrator_params(gamma=m["gamma"]))] = q
            first = False
        if done:
            gone = []
            for mb in e.ensemble_members():
                q = ids.get(mb["id"])
                if q is not None and mb["t"] >= members[q]["t_end"]:
                    res[q] = ({f: e.ensemble_download(mb["id"], f) for f in FIELDS}, mb["cycle"])
                    gone.append(mb["id"])
            e.ensemble_remove(gone)
            for g in gone:
                ids.pop(g)
    return res


def check_same(got, ref, tag):
    for f in FIELDS:
        a, b = np.asarray(got[f]), np.asarray(ref[f])
        assert a.shape == b.shape, (tag, f)
        assert np.array_equal(a, b), (tag, f, np.abs(a - b).max())


def tubes(count, nmin, nmax, t_end, seed):
    ms = problems.tube_ensemble(count, nmin, nmax, t_end=t_end, seed=seed)
    # uneven end times, so members leave while others keep running
    for q, m in enumerate(ms):
        m["t_end"] = t_end * (0.5 + 0.5 * (q % 3) / 2.0)
    return ms


@pytest.mark.parametrize("row_order", ["reference", "search"])
def test_tube_members_match_plain_runs(sphgpu, row_order):
    order = sphgpu.ROW_ORDER_REFERENCE if row_order == "reference" else sphgpu.ROW_ORDER_SEARCH
    members = tubes(5, 60, 240, 0.02, seed=5)
    print(f"[ensemble tubes] {row_order} n={[m['n'] for m in members]}", flush=True)
    hp = sphgpu.hydro_params()
    e = ensemble(sphgpu, 1, order)
    got = run_ensemble(sphgpu, e, members, hp, {}, late=(3,))
    e.close()
    assert sorted(got) == list(range(len(members)))
    for q, m in enumerate(members):
        ref, steps = plain_run(sphgpu, 1, m["state"], m["gamma"], m["t_end"], hp, {}, order)
        fields, cycle = got[q]
        assert cycle == steps, (q, cycle, steps)
        check_same(fields, ref, q)


def test_blast_members_match_plain_runs(sphgpu):
    members = []
    for q, n in enumerate((12, 17, 14)):
        members.append(dict(state=problems.point_blast(n, seed=100 + q), gamma=5.0 / 3.0, t_end=0.004 + 0.002 * q))
    print("[ensemble blasts]", flush=True)
    hp = sphgpu.hydro_params()
    e = ensemble(sphgpu, 2, sphgpu.ROW_ORDER_SEARCH)
    got = run_ensemble(sphgpu, e, members, hp, {}, late=(1,))
    e.close()
    for q, m in enumerate(members):
        ref, steps = plain_run(sphgpu, 2, m["state"], m["gamma"], m["t_end"], hp, {}, sphgpu.ROW_ORDER_SEARCH)
        assert got[q][1] == steps
        check_same(got[q][0], ref, q)


def test_mixed_and_compatible_members(sphgpu):
    """Mixed precision and compatible energy take the same per-member path."""
    members = tubes(3, 80, 160, 0.01, seed=9)
    for prec, kw in ((sphgpu.PRECISION_MIXED, {}), (sphgpu.PRECISION_FP64, dict(compatibleEnergyEvolution=1)),
                     (sphgpu.PRECISION_MIXED, dict(compatibleEnergyEvolution=1))):
        print(f"[ensemble variant] precision={prec} {kw}", flush=True)
        hp = sphgpu.hydro_params(precision=prec, **kw)
        e = ensemble(sphgpu, 1, sphgpu.ROW_ORDER_SEARCH, precision=prec)
        got = run_ensemble(sphgpu, e, members, hp, {})
        e.close()
        for q, m in enumerate(members):
            ref, steps = plain_run(sphgpu, 1, m["state"], m["gamma"], m["t_end"], hp, {}, sphgpu.ROW_ORDER_SEARCH,
                                   precision=prec)
            assert got[q][1] == steps
            check_same(got[q][0], ref, (prec, q))


def test_3d_members_reference_order(sphgpu):
    members = []
    for q, n in enumerate((6, 7)):
        st, _ = problems.converging_flow(n, seed=300 + q)
        members.append(dict(state=st, gamma=5.0 / 3.0, t_end=2.0e-4))
    print("[ensemble 3d]", flush=True)
    hp = sphgpu.hydro_params()
    e = ensemble(sphgpu, 3, sphgpu.ROW_ORDER_REFERENCE)
    got = run_ensemble(sphgpu, e, members, hp, {})
    e.close()
    for q, m in enumerate(members):
        ref, steps = plain_run(sphgpu, 3, m["state"], m["gamma"], m["t_end"], hp, {}, sphgpu.ROW_ORDER_REFERENCE)
        assert got[q][1] == steps
        check_same(got[q][0], ref, q)


def test_serve_results_same_alone_and_in_the_mix(sphgpu):
    """The request loop: arrivals overlap, members join and leave while others
    run; every result equals the request solved alone."""
    reqs = []
    for q, n in enumerate((10, 16, 12, 14)):
        reqs.append(dict(state=problems.point_blast(n, seed=200 + q), gamma=5.0 / 3.0, t_end=0.003 + 0.001 * q,
                         arrival=0.02 * q))
    print("[serve]", flush=True)
    e = ensemble(sphgpu, 2, sphgpu.ROW_ORDER_SEARCH)
    hp = sphgpu.hydro_params()
    out = e.serve(reqs, hp, sphgpu.integrator_params(), max_members=3)
    e.close()
    for q, r in enumerate(reqs):
        o = out[q]
        assert 0.0 <= o["admitted"] <= o["finished"]
        assert o["admitted"] >= r["arrival"] - 1e-9
        ref, steps = plain_run(sphgpu, 2, r["state"], r["gamma"], r["t_end"], hp, {}, sphgpu.ROW_ORDER_SEARCH)
        assert o["steps"] == steps
        got = {"position": o["pos"], "velocity": o["vel"], "mass density": o["rho"],
               "specific thermal energy": o["eps"], "H": o["H"]}
        check_same({k: np.asarray(v).reshape(np.asarray(ref[k]).shape) for k, v in got.items()}, ref, q)


def test_ensemble_errors(sphgpu):
    e = ensemble(sphgpu, 1, sphgpu.ROW_ORDER_SEARCH)
    st = problems.shock_tube(50)
    with pytest.raises(sphgpu.SphGpuError):
        e.set_nodes(10, 10)
    with pytest.raises(sphgpu.SphGpuError):
        e.ensemble_step()                      # no members
    mid = e.ensemble_add(st, 1.4, 1.0e-4)
    while e.ensemble_step() == 0:
        pass
    with pytest.raises(sphgpu.SphGpuError):
        e.ensemble_step()                      # a finished member must leave first
    with pytest.raises(sphgpu.SphGpuError):
        e.ensemble_remove([mid + 7])
    e.ensemble_remove([mid])
    e.close()


def _poisoned_tubes():
    members = tubes(4, 60, 200, 0.015, seed=17)
    bad = {k: np.array(v, dtype=np.float64) for k, v in members[1]["state"].items()}
    bad["eps"][bad["eps"].shape[0] // 3] = np.nan       # a value, not a fault
    members[1] = dict(members[1], state=bad)
    pos = {k: np.array(v, dtype=np.float64) for k, v in members[2]["state"].items()}
    pos["pos"][5, 0] = np.nan
    members[2] = dict(members[2], state=pos)
    return members


@pytest.mark.parametrize("row_order", ["reference", "search"])
def test_failed_members_evicted_others_unchanged(sphgpu, row_order):
    """One tube carries a NaN specific energy, another a NaN position: the
    first fails in the step whose final update turns its state non-finite,
    the second at the first step's neighbour build; both are evicted with
    MEMBER_NONFINITE and every other member equals its solo run bit for bit."""
    order = sphgpu.ROW_ORDER_REFERENCE if row_order == "reference" else sphgpu.ROW_ORDER_SEARCH
    members = _poisoned_tubes()
    print(f"[ensemble failures] {row_order}", flush=True)
    hp = sphgpu.hydro_params()
    e = ensemble(sphgpu, 1, order)
    failed = {}
    got = run_ensemble(sphgpu, e, members, hp, {}, failed=failed)
    e.close()
    assert sorted(failed) == [1, 2], failed
    assert failed[1] == (sphgpu.MEMBER_NONFINITE, 1), failed     # after its first step
    assert failed[2] == (sphgpu.MEMBER_NONFINITE, 0), failed     # before any step
    assert sorted(got) == [0, 3]
    for q in (0, 3):
        m = members[q]
        ref, steps = plain_run(sphgpu, 1, m["state"], m["gamma"], m["t_end"], hp, {}, order)
        assert got[q][1] == steps
        check_same(got[q][0], ref, q)


def test_collapsing_dt_member_evicted(sphgpu):
    """A tube 1e8 times hotter than the others votes a dt near 1e-8; with the
    dt threshold at 1e-6 it fails with MEMBER_DT_COLLAPSE on the first step
    and the others finish as they would alone."""
    members = tubes(3, 60, 120, 0.01, seed=23)
    hot = {k: np.array(v, dtype=np.float64) for k, v in members[0]["state"].items()}
    hot["eps"] *= 1.0e8
    members[0] = dict(members[0], state=hot)
    hp = sphgpu.hydro_params()
    e = ensemble(sphgpu, 1, sphgpu.ROW_ORDER_SEARCH)
    e.ensemble_set_dt_threshold(1.0e-6)
    failed = {}
    got = run_ensemble(sphgpu, e, members, hp, {}, failed=failed)
    e.close()
    assert failed == {0: (sphgpu.MEMBER_DT_COLLAPSE, 1)}, failed
    for q in (1, 2):
        m = members[q]
        ref, steps = plain_run(sphgpu, 1, m["state"], m["gamma"], m["t_end"], hp, {}, sphgpu.ROW_ORDER_SEARCH)
        assert got[q][1] == steps
        check_same(got[q][0], ref, q)


def test_serve_failed_request_isolated(sphgpu):
    """The request loop with a poisoned request: its status says why it
    failed, and every other request equals its lone solve."""
    reqs = []
    for q, n in enumerate((10, 13, 11)):
        reqs.append(dict(state=problems.point_blast(n, seed=400 + q), gamma=5.0 / 3.0, t_end=0.003,
                         arrival=0.01 * q))
    bad = {k: np.array(v, dtype=np.float64) for k, v in reqs[1]["state"].items()}
    bad["eps"][7] = np.inf
    reqs[1] = dict(reqs[1], state=bad)
    e = ensemble(sphgpu, 2, sphgpu.ROW_ORDER_SEARCH)
    hp = sphgpu.hydro_params()
    out = e.serve(reqs, hp, sphgpu.integrator_params(), max_members=3)
    e.close()
    assert out[1]["status"] == sphgpu.MEMBER_NONFINITE
    for q in (0, 2):
        o = out[q]
        assert o["status"] == sphgpu.MEMBER_OK
        r = reqs[q]
        ref, steps = plain_run(sphgpu, 2, r["state"], r["gamma"], r["t_end"], hp, {}, sphgpu.ROW_ORDER_SEARCH)
        assert o["steps"] == steps
        got = {"position": o["pos"], "velocity": o["vel"], "mass density": o["rho"],
               "specific thermal energy": o["eps"], "H": o["H"]}
        check_same({k: np.asarray(v).reshape(np.asarray(ref[k]).shape) for k, v in got.items()}, ref, q)


@pytest.mark.parametrize("field", ["pos", "H"])
def test_plain_sim_rejects_nonfinite_nodes(sphgpu, field):
    """Outside an ensemble a NaN position or H makes the neighbour build fail
    with SPHGPU_ERR_ARG (min/max reductions would pass a NaN over), on the
    single-grid and the multi-level path."""
    for spread in (0.0, 0.6):
        st, _ = problems.random_state(2, 12, seed=31, h_var=spread)
        st = {k: np.array(v, dtype=np.float64) for k, v in st.items()}
        st[field][4, 0] = np.nan
        N = st["pos"].shape[0]
        s = sphgpu.Sim(2)
        s.set_table_bspline(1000)
        s.set_nodes(N, N)
        s.upload_state(st)
        with pytest.raises(sphgpu.SphGpuError, match="non-finite"):
            s.build_neighbors(2.0)
        s.close()


def test_mixed_3d_rejects_positions_beyond_the_fixed_point_range(sphgpu):
    """The 3-D mixed record holds round(x / u) with u = search radius / 2^30:
    a lattice 1e10 from the origin would overflow it and is refused; the FP64
    pass takes it."""
    st, _ = problems.converging_flow(6)
    st = {k: np.array(v, dtype=np.float64) for k, v in st.items()}
    st["pos"] = st["pos"] + 1.0e10
    N = st["pos"].shape[0]
    s = sphgpu.Sim(3)
    s.set_row_order(sphgpu.ROW_ORDER_SEARCH)
    s.set_table_bspline(1000)
    s.set_nodes(N, N)
    s.upload_state(st)
    s.build_neighbors(2.0)
    s.evaluate_derivatives(sphgpu.hydro_params())
    with pytest.raises(sphgpu.SphGpuError, match="fixed-point"):
        s.evaluate_derivatives(sphgpu.hydro_params(precision=sphgpu.PRECISION_MIXED))
    s.close()
