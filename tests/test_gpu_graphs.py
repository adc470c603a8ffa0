"""The device part of a time step replayed from a captured HIP graph
(sphgpu_set_graphs) against the same step launched eagerly: after every
step the dt and the whole state are bitwise equal, for a plain simulation
(1-D tube, 3-D converging flow with varying ghost counts, compatible energy
with grad-h) and for an ensemble whose members join and leave.  Each replay
is checked, with the host's own eager work and synchronisations between
replays."""
import numpy as np
import pytest

from sphgpu import problems

pytestmark = pytest.mark.gpu

FIELDS = ("position", "velocity", "mass density", "specific thermal energy", "H", "pressure", "sound speed",
          "grad h corrections")


def _sim(sg, dim, st, kind, gamma, order, precision, graphs):
    N = st["pos"].shape[0]
    s = sg.Sim(dim)
    s.set_row_order(order)
    s.set_table_bspline(1000)
    s.set_nodes(N, N)
    s.upload_state({k: st[k] for k in ("pos", "vel", "mass", "rho", "eps", "H")})
    s.eos_gamma_law(gamma)
    s.set_boundaries(sg.box_boundaries(dim, kind=kind))
    s.set_precision(precision)
    s.set_graphs(graphs)
    return s


# (name, dim, state, boundaries, gamma, search order, precision, hydro options,
#  steps, replays expected: None where the ghost counts change from step to
#  step (the converging flow, the blast's walls while its smoothing scales
#  settle), so that a key may never repeat and the steps may all run eagerly)
CASES = [
    ("tube1d_fp64_ref", 1, lambda: problems.shock_tube(150), "reflecting", 1.4, 0, 0, {}, 30, True),
    ("tube1d_mixed", 1, lambda: problems.shock_tube(150), "reflecting", 1.4, 1, 1, {}, 30, True),
    ("converging3d_fp64", 3, lambda: problems.converging_flow(7)[0], "reflecting", 5.0 / 3.0, 1, 0, {}, 12, None),
    ("blast2d_mixed", 2, lambda: problems.point_blast(16), "reflecting", 5.0 / 3.0, 1, 1, {}, 20, None),
    ("periodic2d_compat_gradh", 2, lambda: problems.periodic_sinusoid(10, dim=2)[0], "periodic", 1.4, 0, 0,
     dict(compatibleEnergyEvolution=1), 10, True),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_graph_steps_equal_eager_steps(sphgpu, case):
    name, dim, make, kind, gamma, order_flag, prec, hkw, steps, replayed = case
    print(f"[graphs] {name}", flush=True)
    st = make()
    N = st["pos"].shape[0]
    order = sphgpu.ROW_ORDER_SEARCH if order_flag else sphgpu.ROW_ORDER_REFERENCE
    a = _sim(sphgpu, dim, st, kind, gamma, order, prec, graphs=False)
    b = _sim(sphgpu, dim, st, kind, gamma, order, prec, graphs=True)
    hp = sphgpu.hydro_params(precision=prec, **hkw)
    ip = sphgpu.integrator_params(gamma=gamma, gradhCorrection=1 if "gradh" in name else 0)
    for k in range(steps):
        da = a.rk2_step(hp, ip)
        db = b.rk2_step(hp, ip)
        assert da == db, (k, da, db)
        for f in FIELDS:
            assert np.array_equal(a.download(f, rows=N), b.download(f, rows=N)), (k, f)
        assert a.node_counts() == b.node_counts(), k
    captures, replays = b.graph_counts()
    print(f"[graphs] {name}: {captures} captures, {replays} replays", flush=True)
    if replayed:
        assert captures >= 1 and replays >= 1, (captures, replays)
    assert a.graph_counts() == (0, 0)
    a.close()
    b.close()


def test_ensemble_graph_steps_equal_eager_steps(sphgpu):
    """Tubes of different sizes and end times: members leave as they finish
    and one joins late; every member's final state and step count are equal
    with and without graphs, and steps between membership changes replay a
    graph."""
    members = problems.tube_ensemble(5, 60, 200, t_end=0.02, seed=31)
    for q, m in enumerate(members):
        m["t_end"] = 0.02 * (0.5 + 0.25 * q)
    hp = sphgpu.hydro_params(precision=sphgpu.PRECISION_MIXED)
    ip = sphgpu.integrator_params()
    out = []
    for graphs in (False, True):
        e = sphgpu.Sim(1)
        e.set_row_order(sphgpu.ROW_ORDER_SEARCH)
        e.set_table_bspline(1000)
        e.set_precision(sphgpu.PRECISION_MIXED)
        e.set_boundaries(sphgpu.box_boundaries(1))
        e.ensemble_begin()
        e.set_graphs(graphs)
        ids = {}
        for q, m in enumerate(members[:-1]):
            ids[e.ensemble_add(m["state"], m["gamma"], m["t_end"], sphgpu.integrator_params(gamma=m["gamma"]))] = q
        res, steps = {}, 0
        while ids:
            done = e.ensemble_step(hp, ip)
            steps += 1
            if steps == 3:
                m = members[-1]
                ids[e.ensemble_add(m["state"], m["gamma"], m["t_end"],
                                   sphgpu.integrator_params(gamma=m["gamma"]))] = len(members) - 1
            if done:
                gone = []
                for mb in e.ensemble_members():
                    q = ids.get(mb["id"])
                    if q is not None and mb["t"] >= members[q]["t_end"]:
                        res[q] = ({f: e.ensemble_download(mb["id"], f)
                                   for f in ("position", "velocity", "mass density", "specific thermal energy", "H")},
                                  mb["cycle"])
                        gone.append(mb["id"])
                e.ensemble_remove(gone)
                for g in gone:
                    ids.pop(g)
        out.append((res, steps, e.graph_counts()))
        e.close()
    (ra, sa, _), (rb, sb, (cap, rep)) = out
    assert sa == sb and sorted(ra) == sorted(rb) == list(range(len(members)))
    for q in ra:
        assert ra[q][1] == rb[q][1], q
        for f in ra[q][0]:
            assert np.array_equal(ra[q][0][f], rb[q][0][f]), (q, f)
    print(f"[graphs] ensemble: {sb} steps, {cap} captures, {rep} replays", flush=True)
    assert cap >= 1 and rep >= 1, (cap, rep, sb)


def test_autotuned_pair_kernel_same_results(sphgpu):
    """With the autotuner on, a 1-D tube ensemble times every launch variant
    of the pair kernel on its first pass and keeps the fastest; the members'
    results are bitwise those of the untuned run."""
    members = problems.tube_ensemble(4, 100, 400, t_end=0.004, seed=57)
    hp = sphgpu.hydro_params(precision=sphgpu.PRECISION_MIXED)
    ip = sphgpu.integrator_params()
    out = []
    for tune in (False, True):
        sphgpu.autotune(tune)
        e = sphgpu.Sim(1)
        e.set_row_order(sphgpu.ROW_ORDER_SEARCH)
        e.set_table_bspline(1000)
        e.set_precision(sphgpu.PRECISION_MIXED)
        e.set_boundaries(sphgpu.box_boundaries(1))
        e.ensemble_begin()
        ids = [e.ensemble_add(m["state"], m["gamma"], 1.0, sphgpu.integrator_params(gamma=m["gamma"]))
               for m in members]
        for _ in range(6):
            e.ensemble_step(hp, ip)
        out.append({(i, f): e.ensemble_download(i, f) for i in ids
                    for f in ("position", "velocity", "mass density", "specific thermal energy", "H")})
        e.close()
    sphgpu.autotune(False)
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k
    rows = [l.split() for l in sphgpu.autotune_table().splitlines() if l and not l.startswith("#")]
    tuned = [r for r in rows if r[0] == "0" and r[1] == "1" and r[2] == "1" and len(r) == 6 + 8]
    assert tuned, rows
    print("[autotune]", tuned, flush=True)
