// Measurement kernels: a plain streaming copy, used to calibrate HBM traffic
// counters (FETCH_SIZE / WRITE_SIZE per dispatch) against a kernel whose
// traffic is known exactly, and a record gather over the current neighbour
// rows, the ceiling the pair and density kernels are stated against.
#include "internal.hh"
#include "dgeom.hh"
#include "ell.hh"

namespace sphgpu {

namespace {

__global__ void __launch_bounds__(256) k_copy_probe(const double4* __restrict__ src, double4* __restrict__ dst, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

// Gather ceiling of the current neighbour rows: one lane per internal node,
// in the pair and density kernels' work order and XCD-aware block order, walks
// its row (int4 chunks) and loads L 16-B pieces of each neighbour's record from
// an array of L x 16 B per slot, folding the loaded words together and nothing
// more: the time those kernels would take if gathering their records were all
// they did.
template <int L>
__global__ void __launch_bounds__(256) k_gather_probe(const int4* __restrict__ rec, const int32_t* __restrict__ rowCount,
                                                      const int4* __restrict__ nb4, int64_t n, int64_t nInt,
                                                      const int32_t* __restrict__ work, int32_t* __restrict__ out) {
  const int64_t tw = xcdBlock((nInt + 255) / 256) * 256 + threadIdx.x;
  if (tw >= nInt) return;
  const int64_t k = work[tw];
  const int len = rowCount[k];
  int acc = 0;
  for (int c = 0; 4 * c < len; ++c) {
    const int4 j4 = nb4[ellChunk(c, n, k)];
    const int m = len - 4 * c < 4 ? len - 4 * c : 4;
    for (int q = 0; q < m; ++q) {
      const int32_t j = q == 0 ? j4.x : (q == 1 ? j4.y : (q == 2 ? j4.z : j4.w));
#pragma unroll
      for (int l = 0; l < L; ++l) {
        const int4 v = rec[int64_t(j) * L + l];
        acc += v.x ^ v.y ^ v.z ^ v.w;
      }
    }
  }
  out[k] = acc;
}

}  // namespace

double gatherProbe(Sim& s, int loads, int repeats) {
  SG_REQUIRE(s.neighborsValid && s.nSlots > 0 && s.nInt > 0, SPHGPU_ERR_STATE,
             "the gather probe walks the rows of a neighbour build");
  SG_REQUIRE(loads == 2 || loads == 3 || loads == 6, SPHGPU_ERR_ARG, "loads per neighbour must be 2, 3 or 6");
  SG_REQUIRE(repeats >= 1, SPHGPU_ERR_ARG, "repeats must be positive");
  const int64_t n = s.nSlots;
  DBuf rec, out;
  const size_t recBytes = size_t(n) * size_t(loads) * 16;
  rec.ensure(recBytes);
  out.ensure(sizeof(int32_t) * size_t(n));
  SG_CHECK_HIP(hipMemsetAsync(rec.p, 0, recBytes, s.stream));
  const dim3 grid(unsigned(xcdGrid((s.nInt + 255) / 256)));
  auto launch = [&] {
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, grid, dim3(256), 0, s.stream, rec.as<int4>(), s.rowCount.as<int32_t>(),
                         s.nbr.as<int4>(), n, s.nInt, s.work.as<int32_t>(), out.as<int32_t>());
    };
    if (loads == 2) go(k_gather_probe<2>);
    else if (loads == 3) go(k_gather_probe<3>);
    else go(k_gather_probe<6>);
  };
  launch();
  SG_CHECK_HIP(hipEventRecord(s.ev[12], s.stream));
  for (int r = 0; r < repeats; ++r) launch();
  SG_CHECK_HIP(hipGetLastError());
  SG_CHECK_HIP(hipEventRecord(s.ev[13], s.stream));
  SG_CHECK_HIP(hipEventSynchronize(s.ev[13]));
  float ms = 0.f;
  SG_CHECK_HIP(hipEventElapsedTime(&ms, s.ev[12], s.ev[13]));
  rec.release();
  out.release();
  return double(ms) / repeats;
}

double copyProbe(Sim& s, int64_t bytes, int repeats) {
  SG_REQUIRE(bytes >= 32 && bytes % 32 == 0, SPHGPU_ERR_ARG, "copy probe needs a positive multiple of 32 bytes");
  SG_REQUIRE(repeats >= 1, SPHGPU_ERR_ARG, "repeats must be positive");
  DBuf a, b;
  a.ensure(size_t(bytes));
  b.ensure(size_t(bytes));
  const int64_t n = bytes / 32;
  SG_CHECK_HIP(hipMemsetAsync(a.p, 0, size_t(bytes), s.stream));
  const int grid = int(std::min<int64_t>((n + 255) / 256, 256 * 64));
  hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, s.stream, a.as<double4>(), b.as<double4>(), n);
  SG_CHECK_HIP(hipEventRecord(s.ev[12], s.stream));
  for (int r = 0; r < repeats; ++r)
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, s.stream, a.as<double4>(), b.as<double4>(), n);
  SG_CHECK_HIP(hipGetLastError());
  SG_CHECK_HIP(hipEventRecord(s.ev[13], s.stream));
  SG_CHECK_HIP(hipEventSynchronize(s.ev[13]));
  float ms = 0.f;
  SG_CHECK_HIP(hipEventElapsedTime(&ms, s.ev[12], s.ev[13]));
  a.release();
  b.release();
  return double(ms) / repeats;
}

}  // namespace sphgpu
