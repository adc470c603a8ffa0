"""Mixed-precision pass (FP32 pair kinematics and kernel lookups, FP64
positions and per-node accumulators) against the FP64 oracle.

Bar: every derivative field and the summed density within 1e-5 relative
error in L-infinity (max_i |F_i - F_i^ref| / max_i |F_i^ref|) and in L2
(||F - F^ref||_2 / ||F^ref||_2), the neighbour set unchanged, and momentum
conserved to FP32 round-off of the pair terms."""
import numpy as np
import pytest

from sphgpu import problems

pytestmark = pytest.mark.gpu

TOL = 1.0e-5


def cases():
    out = []
    st, _ = problems.periodic_sinusoid(12)
    out.append(("periodic3d", 3, st, "periodic", {}))
    st, _ = problems.converging_flow(10)
    out.append(("converging3d", 3, st, "reflecting", {}))
    st, _ = problems.random_state(3, 9, seed=91, h_var=0.15)
    out.append(("random3d", 3, st, None, {}))
    st, _ = problems.random_state(2, 20, seed=92, h_var=0.15)
    out.append(("random2d", 2, st, None, dict(linearInExpansion=1, quadraticInExpansion=1)))
    st, _ = problems.random_state(1, 150, seed=93, h_var=0.15)
    out.append(("random1d", 1, st, "reflecting", dict(evolveTotalEnergy=1)))
    st, _ = problems.random_state(3, 8, seed=94, h_var=0.1)
    out.append(("noxsph3d", 3, st, None, dict(XSPH=0, correctVelocityGradient=0)))
    # far from the origin, negative coordinates too: the 3-D mixed record's
    # positions are the low 32 bits of round(x / u), here past 2^32 and wrapped
    st, _ = problems.random_state(3, 9, seed=95, h_var=0.15)
    st = dict(st)
    st["pos"] = st["pos"] + np.array([1000.0, -1000.0, 250.5])
    out.append(("shifted3d", 3, st, None, {}))
    return out


CASES = cases()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rel2(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def reference(oracle, dim, st, kind, kw):
    N = st["pos"].shape[0]
    if kind:
        bo = oracle.box_boundaries(dim, kind=kind)
        full, off, nbr, info = oracle.ghost_pipeline(dim, st, bo)
    else:
        full, info, bo = dict(st), None, None
        off, nbr = oracle.connectivity(dim, N, st["pos"], st["H"])
    T = oracle.Table(dim, 1000)
    rho = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr)
    rho_all = np.array(full["rho"], dtype=float)
    rho_all[:N] = rho
    if kind:
        rho_all = oracle.apply_ghosts(dim, N, bo, info["control"], info["sub_of"], rho_all, 0)
    full["rho"] = rho_all
    ref = oracle.eval_derivs(T, oracle.make_params(**kw), N, full, off, nbr)
    return rho, ref, off, nbr


def device(sg, dim, st, kind, kw, precision):
    N = st["pos"].shape[0]
    s = sg.Sim(dim)
    s.set_row_order(sg.ROW_ORDER_SEARCH)
    s.set_table_bspline(1000)
    s.set_nodes(N, N)
    s.upload_state(st)
    if kind:
        s.set_boundaries(sg.box_boundaries(dim, kind=kind))
        s.set_ghost_nodes(2.0)
    else:
        s.build_neighbors(2.0)
    s.set_precision(precision)
    s.sum_density()
    s.evaluate_derivatives(sg.hydro_params(precision=precision, **kw))
    return s


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_mixed_vs_fp64_oracle(sphgpu, oracle, case):
    name, dim, st, kind, kw = case
    print(f"[mixed] {name}", flush=True)
    N = st["pos"].shape[0]
    rho, ref, off, nbr = reference(oracle, dim, st, kind, kw)
    s = device(sphgpu, dim, st, kind, kw, sphgpu.PRECISION_MIXED)
    o2, n2 = s.neighbors()
    assert np.array_equal(o2, off) and np.array_equal(n2, nbr)
    rho_dev = s.download("mass density", rows=N)
    assert rel(rho_dev, rho) <= TOL and rel2(rho_dev, rho) <= TOL
    errs = {}
    for f in sphgpu.DERIV_ALIASES:
        a = s.derivative(f).reshape(ref[f].shape)
        errs[f] = (rel(a, ref[f]), rel2(a, ref[f]))
    print(" ".join(f"{k}={v[0]:.1e}/{v[1]:.1e}" for k, v in errs.items()), flush=True)
    bad = {k: v for k, v in errs.items() if not (v[0] <= TOL and v[1] <= TOL)}
    assert not bad, bad
    # the unordered-pair count is a property of the neighbour set
    fp = device(sphgpu, dim, st, kind, kw, sphgpu.PRECISION_FP64)
    assert s.pair_count() == fp.pair_count()


def test_mixed_momentum_conservation(sphgpu, oracle):
    """Free surface, no ghosts: sum_i m_i DvDt_i vanishes up to the FP32
    rounding of the per-pair products with the partner mass."""
    st, _ = problems.random_state(3, 10, seed=95, h_var=0.1)
    s = device(sphgpu, 3, st, None, {}, sphgpu.PRECISION_MIXED)
    m = st["mass"]
    a = s.derivative("DvDt")
    mom = (m[:, None] * a).sum(0)
    scale = np.abs(m[:, None] * a).sum(0)
    assert np.all(np.abs(mom) <= 1e-6 * scale), (mom, scale)
    e = m * ((st["vel"] * a).sum(1) + s.derivative("DepsDt"))
    escale = np.abs(m * (st["vel"] * a).sum(1)).sum() + np.abs(m * s.derivative("DepsDt")).sum()
    assert abs(e.sum()) <= 1e-6 * escale


def test_mixed_unsupported_cases(sphgpu):
    st, meta = problems.random_state(3, 6, seed=96)
    st["H"] = problems.anisotropic_H(3, meta["n_internal"], 1.51 * meta["dx"])
    N = st["pos"].shape[0]
    s = sphgpu.Sim(3)
    s.set_table_bspline(1000)
    s.set_nodes(N, N)
    s.upload_state(st)
    s.build_neighbors(2.0)
    s.sum_density()
    with pytest.raises(sphgpu.SphGpuError) as e:
        s.evaluate_derivatives(sphgpu.hydro_params(precision=sphgpu.PRECISION_MIXED))
    assert e.value.code == -4
