"""Benchmark of the SPH derivative pass on one MI355X.

One step = Integrator::setGhostNodes (reflecting-box ghosts, neighbour build,
ghost culling) + computeSPHSumMassDensity (with the ghost density refresh) +
SPHHydroBase::evaluateDerivatives over the resident node state, i.e. the
pass the metric "SPH pair-interactions/s (neighbors+derivs)" times.  Workload:
the 3-D converging-flow generator at 128^3 nodes in the unit box with 6
reflecting planes (cold gas, v = -r_hat, +-1% lattice jitter, cubic B-spline
TableKernel, nPerh 1.51, Monaghan-Gingold Q with Cl = Cq = 1).
pair-interactions/s = unordered pairs per pass (the reference's ncalc,
internal-internal and internal-ghost) x steps / wall time.

The same line carries the second half of the metric, "ensemble p50/p99
solve time": 512 batched 1-D shock tubes (200-20000 nodes each, to t = 0.2)
with each tube's time to solution and the one-at-a-time estimate, and the
request loop over 32 Poisson-arriving 2-D point blasts (p50/p99 time to
solution batched and one at a time, error against FP64 solves).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--precision fp64|mixed]
                    [--ensemble 512] [--serve 32] [--periodic 256]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "llnl-spheral_amd"))

import numpy as np  # noqa: E402

METRIC = "SPH pair-interactions/s (neighbors+derivs) at 128³ 3D; ensemble p50/p99 solve time"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
# Floating-point operations of the collapsed isotropic pair body per directed
# neighbour entry (the FP64 kernel in search row order, and the mixed kernel),
# counted term by term from its source (csrc/hydro.hip, k_derivs_iso): a
# reciprocal and a reciprocal square root count as one operation each, an
# FMA as two; compares, selects and conversions count none.  The FP64 kernel
# also executes two Newton steps per reciprocal / reciprocal square root
# (36 FLOPs more per entry).  Mixed: the pair terms and their block sums in
# FP32, and each block of 16 entries adds its 34 partial sums into the FP64
# per-node accumulators (34 / 16 FP64 additions per entry).
PAIR_FLOPS = {"fp64": {"fp64": 174}, "mixed": {"fp32": 174, "fp64": 34 / 16}}
IN_BYTES_PER_NODE = 136      # x(3) v(3) m rho P c omega H(6), FP64
OUT_BYTES_PER_NODE = 576     # the 72 doubles of the derivative fields
NBR_BYTES_PER_ENTRY = 4      # one int32 neighbour index per directed entry
# HBM bytes per pair-kernel dispatch, from calibrated FETCH_SIZE / WRITE_SIZE runs
PROFILES = os.path.join(ROOT, "profiles")
VALU_ISSUE_HZ = 1024 * 2.4e9 / 4     # wavefront instructions per second: 1024 SIMDs, 4 cycles each
# 16-B loads per neighbour of the isotropic pair kernels (96-B FP64 record,
# 48-B 3-D mixed record) and of the density sums (48-B FP64, 32-B mixed
# density record); their gather ceilings are measured on the pass's own rows
GATHER_LOADS = {("fp64", "search"): 6, ("mixed", "search"): 3}
DENSITY_LOADS = {"fp64": 3, "mixed": 2}


def measured_counters(n, precision, row_order, boundaries):
    """Counter data of the pair kernel recorded for this workload (HBM traffic
    per launch, VALU wavefront instructions per launch), or None."""
    key = [n, precision, row_order, boundaries]
    for name in sorted(os.listdir(PROFILES)) if os.path.isdir(PROFILES) else []:
        if not name.endswith("_traffic.json"):
            continue
        with open(os.path.join(PROFILES, name)) as f:
            t = json.load(f)
        if t.get("key") == key:
            return t, os.path.join("profiles", name)
    return None, None


def pair_roofline(precision, row_order, n, boundaries, entries, n_int, n_tot, kernel_ms, gather_floor_ms=None):
    """Roofline of the pair kernel (evaluateDerivatives pair loop + epilogue)
    for one pass: the FLOP roofline from the exact per-entry count (the
    binding one: the larger fraction), the HBM roofline from the algorithmic
    bytes, the VALU issue fraction and the L2 hit rate from the committed
    counters of this workload."""
    alg_bytes = n_int * OUT_BYTES_PER_NODE + n_tot * IN_BYTES_PER_NODE + NBR_BYTES_PER_ENTRY * entries
    t = kernel_ms * 1e-3
    counters, src = measured_counters(n, precision, row_order, boundaries)
    if counters and counters.get("kernel"):
        kernel = counters["kernel"]
    elif precision == "mixed":
        kernel = "k_derivs_iso<float, 3, true, false>"
    else:
        kernel = ("k_derivs_iso<double, 3, true, true>" if row_order == "search"
                  else "k_derivs_fp64<3, true, false, true>")
    hbm = {"achieved": alg_bytes / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": alg_bytes / t / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes": alg_bytes,
           "traffic": counters["traffic_bytes"] if counters else None}
    out = {"kernel": f"{kernel} (evaluateDerivatives pair loop + epilogue)", "traffic_source": src}
    fl = PAIR_FLOPS.get(precision) if row_order == "search" else None
    if fl:
        rates = {k: v * entries / t / 1e12 for k, v in fl.items()}
        peaks = {"fp64": FP64_PEAK_TFLOPS, "fp32": FP32_PEAK_TFLOPS}
        frac = sum(rates[k] / peaks[k] for k in rates)
        total = sum(rates.values())
        out.update({"bound": "valu", "achieved": total, "peak": total / frac if frac else None, "unit": "TFLOP/s",
                    "frac": frac, "traffic": hbm["traffic"],
                    "flops_per_entry": fl, "flops_per_launch": {k: v * entries for k, v in fl.items()},
                    "flop_rate_tflops": rates,
                    "flop_peaks_tflops": {k: peaks[k] for k in rates},
                    "note": "FLOP roofline (vector ALU; no MFMA on a sparse gather): exact FLOPs per directed "
                            "entry x entries / the kernel's event time, against the published FP64 / FP32 peaks; "
                            "peak is their FLOP-weighted combination"})
        out["hbm"] = hbm
    else:
        out.update({"bound": "hbm", "achieved": hbm["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": hbm["frac"], "traffic": hbm["traffic"], "algorithmic_bytes": alg_bytes})
    if counters and counters.get("valu_wave_instructions"):
        issue_ms = 1e3 * counters["valu_wave_instructions"] / VALU_ISSUE_HZ
        out["issue"] = {"valu_wave_instructions": counters["valu_wave_instructions"], "issue_ms": issue_ms,
                        "kernel_ms": kernel_ms, "busy_frac": issue_ms / kernel_ms,
                        "note": "VALU wavefront instructions (SQ_INSTS_VALU) at one per 4 cycles per SIMD, "
                                "1024 SIMDs at 2.4 GHz, over the kernel time"}
    if counters and counters.get("tcc_hit") is not None:
        h, m = counters["tcc_hit"], counters["tcc_miss"]
        out["l2_hit_rate"] = h / (h + m) if h + m else None
    loads = GATHER_LOADS.get((precision, row_order))
    if loads and gather_floor_ms:
        acc = entries * loads
        out["gather"] = {
            "loads_per_neighbour": loads, "loads_per_launch": acc, "achieved_per_s": acc / t,
            "floor_ms": gather_floor_ms, "kernel_ms": kernel_ms, "ceiling_per_s": acc / (gather_floor_ms * 1e-3),
            "frac": gather_floor_ms / kernel_ms,
            "note": "measured ceiling, not a published peak: sphgpu_gather_probe walks the same rows in the "
                    "kernel's work order and only loads its records (loads_per_neighbour 16-B pieces per "
                    "neighbour); frac = that time / the kernel's time"}
    return out


def cpu_info():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(budget_s=10.0, n=48, boundaries="reflecting"):
    """The oracle's OpenMP gather-form restatement, timed on this host on a
    bounded sample (n^3 converging flow, same boundaries as the GPU run)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # test infrastructure, used only as the CPU baseline here
    from sphgpu import problems
    st, meta = problems.converging_flow(n)
    N = meta["n_internal"]
    T = oracle.Table(3, 1000, fast=True)
    p = oracle.make_params()
    threads = oracle.num_threads(fast=True)
    passes, pairs, t_total = 0, 0, 0.0
    t_nb = t_rho = t_dv = 0.0
    bounds = oracle.box_boundaries(3) if boundaries == "reflecting" else []
    while passes == 0 or t_total < budget_s:
        t0 = time.perf_counter()
        if bounds:
            full, off, nbr, info = oracle.ghost_pipeline(3, st, bounds, fast=True)
        else:
            full = st
            off, nbr = oracle.connectivity(3, N, st["pos"], st["H"], 2.0, fast=True)
        t1 = time.perf_counter()
        oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr, gather=True)
        t2 = time.perf_counter()
        oracle.eval_derivs(T, p, N, full, off, nbr, gather=True)
        t3 = time.perf_counter()
        t_nb += t1 - t0
        t_rho += t2 - t1
        t_dv += t3 - t2
        t_total += t3 - t0
        passes += 1
        pairs = int((nbr > np.repeat(np.arange(N), np.diff(off))).sum())   # the reference's ncalc
    model, ncpu = cpu_info()
    return {
        "value": pairs * passes / t_total, "unit": "pair-interactions/s", "cores": threads, "kind": "port",
        "sample": f"converging flow {n}^3 ({N} nodes, {pairs} pairs, boundaries: {boundaries}), {passes} passes "
                  f"of ghosts + neighbour build + sum density + evaluateDerivatives (OpenMP gather form; the "
                  f"ghost generation is the sequential restatement), {t_total:.1f} s",
        "stage_ms": {"neighbors": 1e3 * t_nb / passes, "sum_density": 1e3 * t_rho / passes,
                     "derivatives": 1e3 * t_dv / passes},
        "cpu_model": model, "host_cpus": ncpu,
    }


def oracle_parity(sim, sphgpu, st, n_int, boundaries, row_order):
    """The checker of the measured FP64 pass, untimed: the oracle's FP64
    restatement runs the same pass on the same state on the host (ghosts of
    the reflecting box, exact connectivity, sum density with the ghost
    refresh, evaluateDerivatives) and the device's neighbour CSR and fields
    are compared with it.  CSR byte for byte (rows ascending, the
    ConnectivityMap layout); fields by relative L-inf and L2 against the
    field's largest magnitude / norm, bar 1e-12 in search row order (value for
    value in reference order, 1e-14 for the cube-root fields)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # test infrastructure, the checker here
    t0 = time.perf_counter()
    N = n_int
    bo = oracle.box_boundaries(3, kind="reflecting") if boundaries == "reflecting" else []
    if bo:
        full, off, nbr, info = oracle.ghost_pipeline(3, st, bo)
    else:
        full = {k: np.array(v) for k, v in st.items()}
        off, nbr = oracle.connectivity(3, N, st["pos"], st["H"], 2.0)
        info = None
    o2, n2 = sim.neighbors()
    csr = bool(np.array_equal(o2, off) and np.array_equal(n2, nbr))
    del o2, n2
    T = oracle.Table(3, 1000)
    rho = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr, gather=True)
    rho_all = np.array(full["rho"])
    rho_all[:N] = rho
    if bo:
        rho_all = oracle.apply_ghosts(3, N, bo, info["control"], info["sub_of"], rho_all, 0)
    full["rho"] = rho_all

    def rel(a, b):
        a, b = np.asarray(a, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()
        return [float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)),
                float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))]

    fields = {"rho": rel(sim.download("mass density", rows=N), rho_all[:N])}
    ref = oracle.eval_derivs(T, oracle.make_params(), N, full, off, nbr, gather=True)
    for f in sphgpu.DERIV_ALIASES:
        fields[f] = rel(sim.derivative(f).reshape(ref[f].shape), ref[f])
    worst = max(max(v) for v in fields.values())
    report = {"csr_bitexact": csr, "directed_entries": int(nbr.size), "row_order": row_order,
              "fields_rel_linf_l2": fields, "worst": worst,
              "bar": 1e-12 if row_order == "search" else 0.0,
              "within_bar": csr and (worst <= 1e-12 if row_order == "search" else None),
              "check_s": time.perf_counter() - t0,
              "oracle": "oracle/ FP64 restatement (-ffp-contract=off parity build), the checker; not timed"}
    ref["rho"] = rho_all[:N]
    return report, ref


def compare_to_ref(sim, sphgpu, ref, n_int, bar):
    """The device's density and derivative fields against reference fields
    kept from oracle_parity: relative L-inf and L2 per field."""
    def rel(a, b):
        a, b = np.asarray(a, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()
        return [float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)),
                float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))]
    fields = {"rho": rel(sim.download("mass density", rows=n_int), ref["rho"])}
    for f in sphgpu.DERIV_ALIASES:
        fields[f] = rel(sim.derivative(f).reshape(ref[f].shape), ref[f])
    worst = max(max(v) for v in fields.values())
    return {"fields_rel_linf_l2": fields, "worst": worst, "bar": bar, "within_bar": worst <= bar,
            "oracle": "the FP64 oracle's fields of the same pass (computed once for the FP64 check)"}


def plain_solve(sphgpu, dim, st, gamma, t_end, hp, prec, order, max_steps=None, graphs=False):
    """One problem alone as a plain simulation (sphgpu_rk2_step loop) to its
    end time, or for max_steps steps; returns (wall seconds, steps, final
    fields)."""
    N = st["mass"].shape[0]
    s = sphgpu.Sim(dim)
    s.set_graphs(graphs)
    s.set_row_order(order)
    s.set_table_bspline(1000)
    s.set_precision(prec)
    s.set_nodes(N, N)
    s.upload_state({k: st[k] for k in ("pos", "vel", "mass", "rho", "eps", "H")})
    s.set_boundaries(sphgpu.box_boundaries(dim))
    t0 = time.perf_counter()
    s.eos_gamma_law(gamma)
    ip = sphgpu.integrator_params(gamma=gamma)
    steps = 0
    while s.clock()[0] < t_end and (max_steps is None or steps < max_steps):
        s.rk2_step(hp, ip, max_time=t_end)
        steps += 1
    s.synchronize()
    wall = time.perf_counter() - t0
    out = {f: s.download(f, rows=N) for f in ("velocity", "mass density", "specific thermal energy")}
    s.close()
    return wall, steps, out


def pct(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else None


def ensemble_bench(sphgpu, problems, count, precision, sample, graphs=False):
    """Batched shock tubes (count tubes, 200-20000 nodes, to t = 0.2): every
    member advanced by one launch sequence per step; time-to-solution of each
    tube from the common start.  The one-at-a-time baseline solves `sample`
    tubes spread over the sizes alone and scales their per-step time to all
    tubes by node count (stated as an estimate)."""
    prec = sphgpu.PRECISION_FP64 if precision == "fp64" else sphgpu.PRECISION_MIXED
    order = sphgpu.ROW_ORDER_SEARCH
    members = problems.tube_ensemble(count)
    hp = sphgpu.hydro_params(precision=prec)
    ip = sphgpu.integrator_params()
    e = sphgpu.Sim(1)
    e.set_row_order(order)
    e.set_table_bspline(1000)
    e.set_precision(prec)
    e.set_boundaries(sphgpu.box_boundaries(1))
    e.ensemble_begin()
    e.set_graphs(graphs)
    print(f"[bench] ensemble of {count} tubes, {sum(m['n'] for m in members)} nodes", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ids = {}
    for q, m in enumerate(members):
        ids[e.ensemble_add(m["state"], m["gamma"], m["t_end"], sphgpu.integrator_params(gamma=m["gamma"]))] = q
    rows = {i: members[q]["n"] for i, q in ids.items()}
    # tubes whose final state is checked against an FP64 solve alone: the
    # one-at-a-time sample's tubes of up to 6000 nodes
    by_n = sorted(range(count), key=lambda q: members[q]["n"])
    pick = [by_n[int(round(k * (count - 1) / max(1, sample - 1)))] for k in range(sample)] if sample > 0 else []
    check = {q for q in pick if members[q]["n"] <= 6000} if prec == sphgpu.PRECISION_MIXED else set()
    final = {}
    finish, cycles = {}, {}
    steps = node_steps = 0
    last_report = t0
    while ids:
        done = e.ensemble_step(hp, ip)
        steps += 1
        node_steps += sum(rows[i] for i in ids)
        if done:
            now = time.perf_counter() - t0
            gone = [mb["id"] for mb in e.ensemble_members() if mb["t"] >= members[ids[mb["id"]]]["t_end"]]
            for mb in e.ensemble_members():
                if mb["id"] in gone:
                    cycles[ids[mb["id"]]] = mb["cycle"]
                    if ids[mb["id"]] in check:
                        final[ids[mb["id"]]] = {f: e.ensemble_download(mb["id"], f)
                                                for f in ("mass density", "specific thermal energy", "velocity")}
            e.ensemble_remove(gone)
            for g in gone:
                finish[ids.pop(g)] = now
        if time.perf_counter() - last_report > 20.0:
            last_report = time.perf_counter()
            print(f"[bench] ensemble step {steps}: {len(ids)} tubes left", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    graph_counts = e.graph_counts()
    e.close()
    tts = [finish[q] for q in range(count)]
    # one-at-a-time plain solves of a sample spread over the sizes
    samp = []
    for q in pick:
        m = members[q]
        w, s, _ = plain_solve(sphgpu, 1, m["state"], m["gamma"], m["t_end"], hp, prec, order, max_steps=300,
                              graphs=graphs)
        samp.append((m["n"], w / s, s))
    est = None
    if samp:
        ln = np.log([x[0] for x in samp])
        lt = np.log([x[1] for x in samp])
        per_step = lambda n: float(np.exp(np.interp(np.log(n), ln, lt)))
        est = float(sum(cycles[q] * per_step(members[q]["n"]) for q in range(count)))
    # final states of the checked members against FP64 solves alone
    errs = []
    for q in sorted(final, key=lambda q: members[q]["n"]):
        m = members[q]
        _, st64, ref = plain_solve(sphgpu, 1, m["state"], m["gamma"], m["t_end"], sphgpu.hydro_params(),
                                   sphgpu.PRECISION_FP64, order)
        row = {"n": m["n"], "steps": st64}
        for f, k in (("mass density", "rho"), ("specific thermal energy", "eps"), ("velocity", "vel")):
            a, b = np.asarray(final[q][f]).ravel(), np.asarray(ref[f]).ravel()
            row[k] = [float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)),
                      float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))]
        errs.append(row)
    return {
        "workload": f"{count} planar shock tubes, n log-uniform in [200, 20000] (total {sum(m['n'] for m in members)}), "
                    "rho_L,P_L~U[0.5,2], rho_R,P_R~U[0.05,0.5], gamma in {1.4,5/3}, reflecting walls, t_end 0.2",
        "precision": precision, "hip_graphs": {"on": bool(graphs), "captures": graph_counts[0],
                                               "replays": graph_counts[1]},
        "wall_s": wall, "steps": steps, "node_steps": node_steps, "node_steps_per_s": node_steps / wall,
        "time_to_solution_s": {"p50": pct(tts, 50), "p99": pct(tts, 99), "max": max(tts)},
        "member_steps": {"p50": pct(list(cycles.values()), 50), "max": max(cycles.values())},
        "sequential_plain": {"sample": [{"n": n, "ms_per_step": 1e3 * t, "steps_timed": s} for n, t, s in samp],
                             "estimated_total_s": est,
                             "speedup": (est / wall) if est else None,
                             "method": "per-step time of the sampled tubes alone (first 300 steps), interpolated in "
                                       "log n, times each tube's step count from the ensemble run"},
        "error_vs_fp64": {"per_member": errs,
                          "note": "final states of sampled members (up to 6000 nodes) of the mixed batch against "
                                  "FP64 solves of the same tubes alone: n, steps, [L-inf, L2] per field"},
    }


def serve_bench(sphgpu, problems, count, rate, precision, check, graphs=False):
    """Poisson arrivals of 2-D point blasts through the request loop
    (continuous batching); time-to-solution = finish - arrival."""
    prec = sphgpu.PRECISION_FP64 if precision == "fp64" else sphgpu.PRECISION_MIXED
    order = sphgpu.ROW_ORDER_SEARCH
    reqs = problems.blast_requests(count, rate)
    hp = sphgpu.hydro_params(precision=prec)
    ip = sphgpu.integrator_params()
    print(f"[bench] serving {count} blasts at {rate}/s, {sum(r['n'] for r in reqs)} nodes", file=sys.stderr,
          flush=True)

    def run(max_members, lane_prec=prec):
        e = sphgpu.Sim(2)
        e.set_row_order(order)
        e.set_table_bspline(1000)
        e.set_precision(lane_prec)
        e.set_boundaries(sphgpu.box_boundaries(2))
        e.ensemble_begin()
        e.set_graphs(graphs)
        t0 = time.perf_counter()
        out = e.serve(reqs, sphgpu.hydro_params(precision=lane_prec), ip, max_members=max_members)
        wall = time.perf_counter() - t0
        e.close()
        tts = [o["finished"] - r["arrival"] for o, r in zip(out, reqs)]
        ns = sum(o["steps"] * r["n"] for o, r in zip(out, reqs))
        return out, {"wall_s": wall, "time_to_solution_s": {"p50": pct(tts, 50), "p99": pct(tts, 99),
                                                             "max": max(tts)},
                     "node_steps_per_s": ns / wall, "max_members": max_members}

    out, batched = run(1024)
    print(f"[bench] serving batched: {batched['wall_s']:.1f}s", file=sys.stderr, flush=True)
    _, single = run(1)
    res = {"workload": f"{count} 2-D point blasts (E=1, rho=1, P=1e-6, gamma 5/3, n log-uniform in [1k, 250k], "
                       f"t_end~U[0.05,0.2], reflecting box), Poisson arrivals at {rate}/s",
           "precision": precision, "hip_graphs": bool(graphs), "batched": batched, "one_at_a_time": single}
    if check > 0 and prec == sphgpu.PRECISION_MIXED:
        # error of the mixed-precision results of the batched run against
        # FP64 solves of the same requests alone (the smallest `check`
        # requests; all of them by default): relative L-inf (over the field's
        # largest magnitude) and L2, worst over the requests
        fields = (("mass density", "rho"), ("specific thermal energy", "eps"), ("velocity", "vel"))

        def rel(a, b):
            a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
            return (float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)),
                    float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)))

        linf, l2, per = {}, {}, []
        picked = sorted(range(count), key=lambda q: reqs[q]["n"])[:check]
        refs = {}
        for q in picked:
            r = reqs[q]
            _, steps, ref = plain_solve(sphgpu, 2, r["state"], r["gamma"], r["t_end"], sphgpu.hydro_params(),
                                        sphgpu.PRECISION_FP64, order)
            refs[q] = ref
            row = {"n": r["n"], "steps": steps}
            for f, k in fields:
                li, ll = rel(out[q][k], ref[f])
                linf[k] = max(linf.get(k, 0.0), li)
                l2[k] = max(l2.get(k, 0.0), ll)
                row[k] = [li, ll]
            per.append(row)
        res["error_vs_fp64"] = {"rel_linf": linf, "rel_l2": l2, "requests": len(picked), "bar": 1e-5,
                                "within_bar": all(v <= 1e-5 for v in list(linf.values()) + list(l2.values())),
                                "per_request": per,
                                "note": "final states of the batched mixed run against FP64 solves alone; "
                                        "per_request: n, steps of the FP64 solve, [L-inf, L2] per field"}
        # control: the same FP64 solve with the other row order (only the
        # summation order differs) on the largest checked requests, to show how
        # much a blast amplifies rounding differences by its end time
        ctrl = []
        for q in picked[-2:]:
            r = reqs[q]
            _, _, alt = plain_solve(sphgpu, 2, r["state"], r["gamma"], r["t_end"], sphgpu.hydro_params(),
                                    sphgpu.PRECISION_FP64, sphgpu.ROW_ORDER_REFERENCE)
            ctrl.append({"n": r["n"], **{k: list(rel(alt[f], refs[q][f])) for f, k in fields}})
        res["fp64_row_order_control"] = {"requests": ctrl,
                                         "note": "FP64 reference row order against FP64 search row order, "
                                                 "[L-inf, L2] per field at the end time"}
        # the same requests served in FP64: the lane for callers that need the
        # 1e-5 bar on the longest blasts, checked against the same solves alone
        out64, lane = run(1024, sphgpu.PRECISION_FP64)
        worst = 0.0
        for q in picked:
            for f, k in fields:
                worst = max(worst, rel(out64[q][k], refs[q][f])[0])
        lane["error_vs_fp64_alone"] = {"max_rel_linf": worst, "requests": len(picked)}
        lane["p99_vs_mixed"] = lane["time_to_solution_s"]["p99"] / batched["time_to_solution_s"]["p99"]
        res["fp64_lane"] = lane
        print(f"[bench] serving FP64 lane: {lane['wall_s']:.1f}s, worst {worst:.2g}", file=sys.stderr, flush=True)
    return res


def periodic_bench(sphgpu, problems, n, steps):
    """Periodic box (density 1 + 0.1 sin sin sin, Gaussian random velocity
    sigma 0.1 with a k^-2 spectrum, gamma 1.4, u = 1): mixed precision against
    the FP64 path, after one pass (sum density + evaluateDerivatives on the
    same state) and after `steps` SynchronousRK2 steps, plus the drift of
    total momentum and energy of both."""
    st, meta = problems.periodic_sinusoid(n)
    N = st["mass"].shape[0]
    print(f"[bench] periodic {n}^3: {N} nodes, {steps} steps", file=sys.stderr, flush=True)
    m = st["mass"]

    def totals(v, eps):
        p = (m[:, None] * v).sum(0)
        return p, float((0.5 * m * (v * v).sum(1) + m * eps).sum())

    def rel(a, b):
        a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
        scale = max(np.abs(b).max(), 1e-300)
        return {"linf": float(np.abs(a - b).max() / scale),
                "l2": float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300 * np.sqrt(a.size)))}

    out = {}
    for prec_name in ("fp64", "mixed"):
        prec = sphgpu.PRECISION_FP64 if prec_name == "fp64" else sphgpu.PRECISION_MIXED
        s = sphgpu.Sim(3)
        s.set_row_order(sphgpu.ROW_ORDER_SEARCH)
        s.set_table_bspline(1000)
        s.set_precision(prec)
        s.set_nodes(N, N)
        s.upload_state({k: st[k] for k in ("pos", "vel", "mass", "rho", "eps", "H")})
        s.set_boundaries(sphgpu.box_boundaries(3, kind="periodic"))
        s.eos_gamma_law(1.4)
        hp = sphgpu.hydro_params(precision=prec)
        ip = sphgpu.integrator_params(gamma=1.4)
        # one pass on the initial state
        s.set_ghost_nodes(2.0)
        s.sum_density()
        s.evaluate_derivatives(hp)
        pas = {"rho": s.download("mass density", rows=N), "DvDt": s.derivative("DvDt"),
               "DepsDt": s.derivative("DepsDt")}
        s.upload("mass density", st["rho"])
        s.eos_gamma_law(1.4)
        p0, e0 = totals(st["vel"], st["eps"])
        t0 = time.perf_counter()
        drift = []
        for k in range(steps):
            s.rk2_step(hp, ip)
            if (k + 1) % max(1, steps // 10) == 0 or k + 1 == steps:
                v = s.download("velocity", rows=N).reshape(N, 3)
                eps = s.download("specific thermal energy", rows=N)
                p, e = totals(v, eps)
                drift.append({"step": k + 1, "momentum": [float(x) for x in (p - p0)], "energy_rel": (e - e0) / e0})
                print(f"[bench] periodic {prec_name} step {k + 1}", file=sys.stderr, flush=True)
        s.synchronize()
        wall = time.perf_counter() - t0
        fin = {"rho": s.download("mass density", rows=N), "eps": s.download("specific thermal energy", rows=N),
               "vel": s.download("velocity", rows=N)}
        mom_scale = float((m[:, None] * np.abs(st["vel"])).sum())
        out[prec_name] = {"pass": pas, "final": fin, "wall_s": wall, "drift": drift, "ms_per_step": 1e3 * wall / steps,
                          "momentum_scale": mom_scale}
        s.close()
    res = {"workload": f"periodic box {n}^3 ({N} nodes): density 1+0.1 sin(2pi x)sin(2pi y)sin(2pi z), Gaussian random "
                       "velocity sigma 0.1 (k^-2), gamma 1.4, u 1; 3 PeriodicBoundary pairs; search row order",
           "steps": steps,
           "pass_error_mixed_vs_fp64": {k: rel(out["mixed"]["pass"][k], out["fp64"]["pass"][k])
                                        for k in ("rho", "DvDt", "DepsDt")},
           "final_error_mixed_vs_fp64": {k: rel(out["mixed"]["final"][k], out["fp64"]["final"][k])
                                         for k in ("rho", "eps", "vel")}}
    for p in ("fp64", "mixed"):
        res[p] = {"ms_per_step": out[p]["ms_per_step"], "wall_s": out[p]["wall_s"],
                  "momentum_scale": out[p]["momentum_scale"], "drift": out[p]["drift"]}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--row-order", choices=["search", "reference"], default="search",
                    help="search: device search order; reference: rows by node index (bitwise reference sums)")
    ap.add_argument("--boundaries", choices=["reflecting", "none"], default="reflecting",
                    help="reflecting: 6 ReflectingBoundary planes on the unit box, ghosts made on the device "
                         "every step (the benchmark configuration); none: free surface")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed-pass", type=int, default=1,
                    help="1: also time the same pass in mixed precision (reported under mixed_pass)")
    ap.add_argument("--rk2-steps", type=int, default=5,
                    help="SynchronousRK2 steps timed after the pass measurement (reported under rk2; 0 = off)")
    ap.add_argument("--probe-mib", type=int, default=256,
                    help="size of the plain copy kernel run after the timed steps (counter calibration); 0 = off")
    ap.add_argument("--ensemble", type=int, default=512,
                    help="batched shock-tube ensemble with this many tubes (512 = the benchmark; 0 = off)")
    ap.add_argument("--ensemble-sample", type=int, default=8,
                    help="tubes solved alone for the one-at-a-time estimate")
    ap.add_argument("--serve", type=int, default=32,
                    help="serving loop with this many blast requests (0 = off)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-check", type=int, default=1 << 30,
                    help="requests (the smallest first) re-solved alone in FP64 for the error; default all")
    ap.add_argument("--batch-precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: launch shapes of the 1-D/2-D pair kernel from the committed tuning table "
                         "(llnl-spheral_amd/tuning/), timing any workload class it lacks")
    ap.add_argument("--graphs", type=int, default=1,
                    help="1: ensemble, serving and their lone baselines replay the device part of each step from a "
                         "captured HIP graph")
    ap.add_argument("--only-batch", action="store_true", help="skip the 128^3 pass measurement")
    ap.add_argument("--periodic", type=int, default=256,
                    help="periodic mixed-vs-FP64 comparison at this lattice size (256 = the configuration; 0 = off)")
    ap.add_argument("--periodic-steps", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the pass would shard by spatial domain with one halo exchange per step).", file=sys.stderr)
        sys.exit(2)

    import sphgpu
    from sphgpu import problems

    tuning = None
    if args.autotune:
        tuning = {"table_file": None, "imported": 0}
        path = os.path.join(ROOT, "llnl-spheral_amd", "tuning", "mi355x_gfx950.txt")
        if os.path.exists(path):
            tuning["table_file"] = os.path.relpath(path, ROOT)
            tuning["imported"] = sphgpu.autotune_load(open(path).read())
        sphgpu.autotune(True)

    def tuning_report():
        if tuning is None:
            return None
        rows = [l.split() for l in sphgpu.autotune_table().splitlines() if l and not l.startswith("#")]
        return dict(tuning, decisions=[{"kernel": "k_derivs_iso", "dim": int(r[1]),
                                        "precision": "mixed" if r[2] == "1" else "fp64",
                                        "row_bucket": int(r[3]), "size_bucket": int(r[4]), "variant": int(r[5]),
                                        "ms": [float(x) for x in r[6:]]} for r in rows if r[0] == "0"],
                    variants="v = 2 * log2(workgroup / 64) + (table in LDS); 4 variants when the table does not fit")

    if args.only_batch:
        line = {"metric": METRIC, "n_gpus": 1}
        if args.periodic:
            line["periodic"] = periodic_bench(sphgpu, problems, args.periodic, args.periodic_steps)
        if args.ensemble:
            line["ensemble"] = ensemble_bench(sphgpu, problems, args.ensemble, args.batch_precision,
                                              args.ensemble_sample, graphs=args.graphs)
        if args.serve:
            line["serving"] = serve_bench(sphgpu, problems, args.serve, args.serve_rate, args.batch_precision,
                                          args.serve_check, graphs=args.graphs)
        line["autotune"] = tuning_report()
        print(json.dumps(line), flush=True)
        return

    t_setup = time.perf_counter()
    st, meta = problems.converging_flow(args.n)
    N = meta["n_internal"]
    sim = sphgpu.Sim(3)
    sim.set_row_order(sphgpu.ROW_ORDER_SEARCH if args.row_order == "search" else sphgpu.ROW_ORDER_REFERENCE)
    sim.set_table_bspline(1000)
    sim.set_nodes(N, N)
    sim.upload_state(st)
    ghosts = args.boundaries == "reflecting"
    if ghosts:
        sim.set_boundaries(sphgpu.box_boundaries(3))
    prec = sphgpu.PRECISION_FP64 if args.precision == "fp64" else sphgpu.PRECISION_MIXED
    params = sphgpu.hydro_params(precision=prec)
    sim.set_precision(prec)
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f}s N={N}", file=sys.stderr, flush=True)

    def measure(params_now):
        def step():
            if ghosts:
                sim.set_ghost_nodes(2.0, cull=True)   # ghosts + neighbour build + culling
            else:
                sim.build_neighbors(2.0)
            sim.sum_density()
            sim.evaluate_derivatives(params_now)

        for _ in range(args.warmup):
            step()
        sim.synchronize()
        stage = np.zeros(6)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
            t = sim.timings()
            stage += [t["neighbors_ms"], t["density_ms"], t["derivs_ms"], t["pack_ms"],
                      t["ghosts_ms"] if ghosts else 0.0, t["cull_ms"] if ghosts else 0.0]
        sim.synchronize()
        wall = time.perf_counter() - t0
        stage /= args.steps
        return wall, stage

    def stage_dict(stage):
        return {"ghosts": stage[4], "neighbors": stage[0], "cull": stage[5], "sum_density": stage[1],
                "derivatives": stage[2], "pack": stage[3]}

    wall, stage = measure(params)

    def gather_floors(pname):
        # the gather ceilings of the pair kernel and the density sum on this
        # pass's rows (the probe walks the rows of the last build)
        pl = GATHER_LOADS.get((pname, args.row_order))
        dl = DENSITY_LOADS[pname]
        return {"pair_ms": sim.gather_probe(pl) if pl else None, "pair_loads": pl,
                "density_ms": sim.gather_probe(dl), "density_loads": dl}

    floors = gather_floors(args.precision)
    parity, oracle_ref = None, None
    if not args.no_cpu and args.precision == "fp64":
        try:
            parity, oracle_ref = oracle_parity(sim, sphgpu, st, N, args.boundaries, args.row_order)
        except Exception as e:  # the check is reported, never required
            parity = {"error": str(e)}
        print(f"[bench] oracle check: {parity.get('check_s', 0):.1f}s", file=sys.stderr, flush=True)
    pairs = sim.pair_count()
    entries = sim.neighbor_count()
    ms_step = 1e3 * wall / args.steps
    n_tot = sim.node_counts()[1]
    line = {
        "metric": METRIC,
        "value": pairs * args.steps / wall,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"converging flow {args.n}^3 3D (cold gas, v=-r_hat, 1% jitter), "
                        + ("reflecting box: 6 ReflectingBoundary planes, ghosts rebuilt on the device every step"
                           if ghosts else "free surface (no boundaries)"),
            "nodes": N, "ghost_nodes": n_tot - N, "pairs": pairs, "directed_entries": entries,
            "kernel": "TableKernel(BSpline,1000)", "nPerh": 1.51, "Q": "MonaghanGingold Cl=Cq=1",
            "options": "XSPH, correctVelocityGradient, no grad-h, no compatible energy",
            "precision": args.precision,
            "row_order": args.row_order,
            "stage_ms": stage_dict(stage),
        },
        "roofline": pair_roofline(args.precision, args.row_order, args.n, args.boundaries, entries, N, n_tot,
                                  stage[2], floors["pair_ms"]),
        "density_gather": {"loads_per_neighbour": floors["density_loads"], "floor_ms": floors["density_ms"],
                           "stage_ms": stage[1], "frac": floors["density_ms"] / stage[1],
                           "note": "sum density stage (the gather kernel plus the ghost density refresh) "
                                   "against the measured gather ceiling of its records on the same rows"},
    }
    if parity is not None:
        line["parity_vs_oracle"] = parity
    if args.precision == "fp64" and args.mixed_pass and args.row_order == "search":
        # the same pass on the same state in mixed precision (FP32 pair
        # arithmetic, FP64 positions and sums)
        sim.set_precision(sphgpu.PRECISION_MIXED)
        wall_m, stage_m = measure(sphgpu.hydro_params(precision=sphgpu.PRECISION_MIXED))
        floors_m = gather_floors("mixed")
        line["mixed_pass"] = {
            "value": sim.pair_count() * args.steps / wall_m, "unit": "pair-interactions/s",
            "ms_per_step": 1e3 * wall_m / args.steps, "steps": args.steps, "warmup": args.warmup,
            "speedup_vs_fp64": wall / wall_m, "dtype": "f32+f64", "stage_ms": stage_dict(stage_m),
            "roofline": pair_roofline("mixed", "search", args.n, args.boundaries, sim.neighbor_count(), N,
                                      sim.node_counts()[1], stage_m[2], floors_m["pair_ms"]),
            "density_gather": {"loads_per_neighbour": floors_m["density_loads"], "floor_ms": floors_m["density_ms"],
                               "stage_ms": stage_m[1], "frac": floors_m["density_ms"] / stage_m[1]},
        }
        if oracle_ref is not None:
            line["mixed_pass"]["parity_vs_oracle"] = compare_to_ref(sim, sphgpu, oracle_ref, N, 1e-5)
        oracle_ref = None
        sim.set_precision(prec)
        sim.sum_density()
    if args.rk2_steps > 0:
        # full time steps on the same resident state: ghosts + neighbours, dt
        # vote, two derivative passes, updates, sum density, enforcement
        hp = sphgpu.hydro_params(precision=prec)
        ip = sphgpu.integrator_params(gamma=meta["gamma"])
        sim.eos_gamma_law(meta["gamma"])
        sim.rk2_step(hp, ip)                      # warm-up step
        sim.synchronize()
        t0 = time.perf_counter()
        dts = [sim.rk2_step(hp, ip) for _ in range(args.rk2_steps)]
        sim.synchronize()
        wall = time.perf_counter() - t0
        line["rk2"] = {"steps": args.rk2_steps, "ms_per_step": 1e3 * wall / args.rk2_steps, "dt": dts,
                       "step": "SynchronousRK2: ghosts + neighbours, dt vote, 2 evaluateDerivatives, updates, "
                               "sum density, boundary enforcement",
                       "pairs_per_s": 2 * pairs * args.rk2_steps / wall}
    if args.probe_mib > 0:
        nb = args.probe_mib << 20
        ms = sim.copy_probe(nb, 5)
        line["copy_probe"] = {"bytes_read": nb, "bytes_written": nb, "ms": ms,
                              "GBps": 2 * nb / (ms * 1e-3) / 1e9,
                              "kernel": "k_copy_probe (double4 streaming copy)"}
    # the ensemble and serving configurations; a failure here is reported in
    # the line and does not drop the pass measurement
    if args.ensemble:
        try:
            line["ensemble"] = ensemble_bench(sphgpu, problems, args.ensemble, args.batch_precision,
                                              args.ensemble_sample, graphs=args.graphs)
        except Exception as e:
            line["ensemble"] = {"error": str(e)}
    if args.serve:
        try:
            line["serving"] = serve_bench(sphgpu, problems, args.serve, args.serve_rate, args.batch_precision,
                                          args.serve_check, graphs=args.graphs)
        except Exception as e:
            line["serving"] = {"error": str(e)}
    if args.periodic:
        sim.close()       # the 256^3 periodic box needs the memory
        try:
            line["periodic"] = periodic_bench(sphgpu, problems, args.periodic, args.periodic_steps)
        except Exception as e:
            line["periodic"] = {"error": str(e)}
    line["autotune"] = tuning_report()
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(args.cpu_budget, boundaries=args.boundaries)
        except Exception as e:  # the baseline is reported, never required
            line["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
