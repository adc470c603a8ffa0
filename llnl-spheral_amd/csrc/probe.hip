// Plain streaming copy, used to calibrate HBM traffic counters (FETCH_SIZE /
// WRITE_SIZE per dispatch) against a kernel whose traffic is known exactly.
#include "internal.hh"

namespace sphgpu {

namespace {

__global__ void __launch_bounds__(256) k_copy_probe(const double4* __restrict__ src, double4* __restrict__ dst, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

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
