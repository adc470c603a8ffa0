// Blocked-transposed ELL layout of the neighbour rows.
//
// Entry w of the row of sorted slot k lives at int index ((w/4)*n + k)*4 + w%4:
// four consecutive entries of a row form one int4, and the int4s of consecutive
// slots are adjacent, so a wavefront of consecutive slots reads or writes one
// contiguous 1 KiB block per chunk.  Capacities are multiples of 4.
#pragma once
#include <cstddef>
#include <cstdint>
#include <hip/hip_runtime.h>

namespace sphgpu {

__host__ __device__ __forceinline__ int64_t ellChunk(int64_t c, int64_t n, int64_t k) { return c * n + k; }
__host__ __device__ __forceinline__ int64_t ellIndex(int64_t w, int64_t n, int64_t k) {
  return ((w >> 2) * n + k) * 4 + (w & 3);
}
inline size_t ellBytes(int64_t n, int cap) { return sizeof(int32_t) * size_t(n > 0 ? n : 1) * size_t(cap); }

}  // namespace sphgpu
