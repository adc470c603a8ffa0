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


def run_ensemble(sg, e, members, hp, ip_kw, late=()):
    """Step until every member finished; members in `late` join after the
    first step.  Returns {index: (fields, cycle)}."""
    ip = sg.integrator_params(**ip_kw)
    ids = {}
    for q, m in enumerate(members):
        if q not in late:
            ids[e.ensemble_add(m["state"], m["gamma"], m["t_end"], sg.integrator_params(gamma=m["gamma"]))] = q
    res = {}
    first = True
    while ids:
        done = e.ensemble_step(hp, ip)
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
    for prec, kw in ((sphgpu.PRECISION_MIXED, {}), (sphgpu.PRECISION_FP64, dict(compatibleEnergyEvolution=1))):
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
