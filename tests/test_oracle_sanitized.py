"""The oracle's C++ under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle sanitize` builds the parity library with both sanitizers; a
child Python with the sanitizer runtimes preloaded runs the oracle's entry
points on small cases of every kind the parity tests use (1-D/2-D/3-D,
reflecting and periodic ghosts with culling, scatter and gather forms, the
option variants, compatible energy, grad-h, ideal H, the gamma law).  Any
out-of-bounds access, use after free or undefined behaviour ends the child
with a sanitizer report and fails the test.  Host code only."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

SCRIPT = r"""
import sys
sys.path[:0] = [%(pkg)r, %(oracle)r]
import numpy as np
import oracle
from sphgpu import problems
for dim, n in ((1, 40), (2, 9), (3, 5)):
    T = oracle.Table(dim, 1000)
    st, _ = problems.random_state(dim, n, seed=3, jitter=0.2, h_var=0.2)
    N = st["pos"].shape[0]
    for kind in ("reflecting", "periodic"):
        full, off, nbr, info = oracle.ghost_pipeline(dim, st, oracle.box_boundaries(dim, kind=kind), 2.0, True)
        rho = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr)
        rho_g = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr, gather=True)
        assert np.array_equal(rho, rho_g)
        for kw in ({}, dict(XSPH=0), dict(correctVelocityGradient=0), dict(epsTensile=0.3),
                   dict(linearInExpansion=1, quadraticInExpansion=1)):
            p = oracle.make_params(**kw)
            a = oracle.eval_derivs(T, p, N, full, off, nbr)
            b = oracle.eval_derivs(T, p, N, full, off, nbr, gather=True)
            for f in oracle.OUTPUT_FIELDS:
                assert np.array_equal(a[f], b[f]), (dim, kind, kw, f)
        d = oracle.eval_derivs(T, oracle.make_params(), N, full, off, nbr, pair_acc=True)
        om = oracle.omega_gradh(T, N, full["pos"], full["H"], off, nbr)
        assert np.all(np.isfinite(om))
    Hs = np.array(full["H"], dtype=np.float64).copy()
    oracle.ideal_h_pass(T, oracle.make_params(), N, full["pos"], Hs, np.full(N, 1.0), off, nbr, 1e-10)
    P, c = oracle.gamma_law(st["rho"], st["eps"], 1.4)
import integrator
st = problems.shock_tube(60)
r = integrator.RK2(1, dict(st), oracle.box_boundaries(1), 1.4, compatible=True)
for _ in range(3):
    r.step()
st2, _ = problems.random_state(2, 8, seed=5, vel_scale=0.3)
r = integrator.RK2(2, {k: st2[k] for k in ("pos", "vel", "mass", "rho", "eps", "H")},
                   oracle.box_boundaries(2, kind="periodic"), 1.4, gradh=True)
for _ in range(2):
    r.step()
print("sanitized oracle ok")
"""


def test_oracle_under_asan_and_ubsan():
    try:
        asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                              check=True).stdout.strip()
        ubsan = subprocess.run(["gcc", "-print-file-name=libubsan.so"], capture_output=True, text=True,
                               check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        pytest.skip("no gcc sanitizer runtimes")
    if not (os.path.isabs(asan) and os.path.exists(asan) and os.path.isabs(ubsan) and os.path.exists(ubsan)):
        pytest.skip("no gcc sanitizer runtimes")
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=f"{asan}:{ubsan}", SPHGPU_ORACLE_LIB=os.path.join(ORACLE, "_build", "liboracle_san.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               OMP_NUM_THREADS="2")
    code = SCRIPT % {"pkg": os.path.join(ROOT, "llnl-spheral_amd"), "oracle": ORACLE}
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    report = r.stdout[-3000:] + r.stderr[-6000:]
    assert r.returncode == 0, report
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
    assert "sanitized oracle ok" in r.stdout, report
