"""sphgpu_conservation_noalloc, the allocation-free entry point, against
sphgpu_conservation on the same rows: bitwise equal, on two calls in a row
(the second over fewer rows, so stale partial sums of the first would show),
for a plain simulation and an ensemble member's rows; a row range outside the
internal rows is rejected."""
import ctypes

import numpy as np
import pytest

from sphgpu import problems

pytestmark = pytest.mark.gpu


def _sim(sg, dim, n, seed):
    st, _ = problems.random_state(dim, n, seed=seed, vel_scale=0.4)
    N = st["pos"].shape[0]
    s = sg.Sim(dim)
    s.set_table_bspline(1000)
    s.set_nodes(N, N)
    s.upload_state(st)
    return s, N


def _hip():
    """The HIP runtime the library itself runs on (the test, as the caller,
    allocates the output buffer)."""
    import sphgpu
    sphgpu.lib()
    for line in open("/proc/self/maps"):
        if "libamdhip64" in line:
            h = ctypes.CDLL(line.split()[-1])
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            return h
    raise RuntimeError("libamdhip64 is not loaded")


class DeviceOut:
    """Nine doubles of device memory, filled with NaN before each call."""

    def __init__(self):
        self.h = _hip()
        self.p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(self.p), 9 * 8) == 0

    def fill_nan(self):
        a = np.full(9, np.nan)
        assert self.h.hipMemcpy(self.p, a.ctypes.data, 72, 1) == 0      # host to device

    def read(self):
        a = np.zeros(9)
        assert self.h.hipMemcpy(a.ctypes.data, self.p, 72, 2) == 0      # device to host
        return a

    def close(self):
        self.h.hipFree(self.p)


def _noalloc(s, first, rows, out):
    out.fill_nan()
    s.conservation_noalloc(first, rows, out.p.value)
    s.synchronize()
    return out.read()


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_noalloc_totals_equal_the_allocating_call(sphgpu, dim):
    print(f"[noalloc] dim={dim}", flush=True)
    big, nb = _sim(sphgpu, dim, {1: 400000, 2: 700, 3: 80}[dim], seed=dim)    # > 256 x 256 rows: capped partials
    small, ns = _sim(sphgpu, dim, {1: 300, 2: 20, 3: 7}[dim], seed=10 + dim)
    out = DeviceOut()
    a = _noalloc(big, 0, nb, out)
    assert np.array_equal(a, big.conservation())
    b = _noalloc(small, 0, ns, out)
    assert np.array_equal(b, small.conservation())
    c = _noalloc(big, nb // 3, nb // 5, out)                  # a row range inside the internal rows
    assert np.isfinite(c).all() and c[0] > 0
    out.close()
    big.close()
    small.close()


def test_noalloc_ensemble_member_rows(sphgpu):
    e = sphgpu.Sim(1)
    e.set_table_bspline(1000)
    e.set_boundaries(sphgpu.box_boundaries(1, kind="reflecting"))
    e.ensemble_begin()
    ids = [e.ensemble_add(problems.shock_tube(n), 1.4, 0.01) for n in (500, 120)]
    e.ensemble_step()
    out = DeviceOut()
    for mid in ids + ids[::-1]:
        mb = next(m for m in e.ensemble_members() if m["id"] == mid)
        got = _noalloc(e, mb["first"], mb["rows"], out)
        assert np.array_equal(got, e.conservation(member=mid)), mid
    out.close()
    e.close()


def test_noalloc_rejects_rows_outside_the_internal_nodes(sphgpu):
    s, n = _sim(sphgpu, 2, 10, seed=3)
    out = DeviceOut()
    for first, rows in ((0, n + 1), (-1, 5), (n - 2, 3), (0, 0)):
        with pytest.raises(sphgpu.SphGpuError, match="row range"):
            s.conservation_noalloc(first, rows, out.p.value)
    out.close()
    s.close()
