// Device ConnectivityMap build.
//
// Contract (src/Neighbor/ConnectivityMap.cc:630-749): for an internal node i, j
// is a neighbour iff |H_i r_ij|^2 <= ext^2 or |H_j r_ij|^2 <= ext^2 (FP64,
// r_ij = r_i - r_j, SymTensor.Vector and magnitude2 summed left to right), j != i.
// The reference orders each row by node index (:735-738); that order is kept on
// request (SPHGPU_ROW_ORDER_REFERENCE, needed for reference-order FP64 sums),
// otherwise rows keep the deterministic search order.  How candidates are found
// is free (the reference's NestedGridNeighbor only has to return a superset):
//
//  1. bounds: min/max position, the smallest Gershgorin lower bound of
//     lambda_min(H) over all nodes, and whether every H is h I.  |H r| <= ext
//     implies |r| <= R = ext/min(lambda) (plus a 1e-6 relative margin).
//  2. counting sort of nodes by cell.  Cells are numbered brick by brick
//     (4^nu cells per brick, Z-order inside) so a brick's nodes are contiguous;
//     each cell is then ordered by node index (the order depends only on
//     positions and indices).  Records of all nodes are packed in this order.
//  3. rows, one thread per internal node, into a blocked-transposed ELL: entry
//     w of slot k at nbr[((w/4)*n + k)*4 + w%4], so four entries are one int4
//     and a wavefront's int4s are contiguous (coalesced writes here, coalesced
//     int4 reads in the passes):
//      - 3-D with isotropic H: cells of edge R/2, one workgroup per brick
//        stages the 8^3 cells around it in LDS (FP32 offsets from the brick
//        origin), each node scans its 5^3 cells; the FP32 estimate of h^2|r|^2
//        decides every pair farther than 1e-5 (relative) from the threshold,
//        the rest take the exact FP64 reference test;
//      - otherwise: cells of edge R, 3^nu stencil read from global memory.
//     Rows longer than the ELL capacity are counted, the build grows the
//     capacity and repeats the row pass.
//  4. optionally, rows sorted by node index (one wavefront per row, bitonic
//     sort in registers).
// Ensembles: bounds and grid per member (each member's grid is the one it
// would get alone); a member's cells are numbered after those of the members
// before it, so the sort groups slots by member, and a row only looks at its
// own member's cells.
#include "internal.hh"
#include "dgeom.hh"
#include "records.hh"
#include "ell.hh"
#include <cstring>

namespace sphgpu {

namespace {

constexpr int kBlock = 256;

inline int nblocks(int64_t n, int bs = kBlock) { return int(std::max<int64_t>(1, (n + bs - 1) / bs)); }

// Gershgorin lower bound of the smallest eigenvalue of a SymTensor.
template <int D> __device__ double gershgorinMin(const double* h) {
  if constexpr (D == 1) return h[0];
  else if constexpr (D == 2) return dmin(h[0] - fabs(h[1]), h[2] - fabs(h[1]));
  else return dmin(h[0] - fabs(h[1]) - fabs(h[2]), dmin(h[3] - fabs(h[1]) - fabs(h[4]), h[5] - fabs(h[2]) - fabs(h[4])));
}
template <int D> __device__ bool isIsotropic(const double* h) {
  if constexpr (D == 1) return true;
  else if constexpr (D == 2) return h[1] == 0.0 && h[2] == h[0];
  else return h[1] == 0.0 && h[2] == 0.0 && h[4] == 0.0 && h[3] == h[0] && h[5] == h[0];
}

// reduced values: lo[3] hi[3] min Gershgorin bound, isotropic flag, max diagonal
// of H, max Gershgorin bound, non-finite flag (1 when a position or H entry is
// NaN or infinite: min and max would pass a NaN over)
constexpr int kNB = 11;
constexpr int kNonFinite = 10;
__host__ __device__ __forceinline__ bool isMaxSlot(int k) { return (k >= 3 && k < 6) || k >= 8; }
#define SG_BOUNDS_IDENT {1e300, 1e300, 1e300, -1e300, -1e300, -1e300, 1e300, 1.0, -1e300, -1e300, 0.0}

// Bounds a grid can be made from: finite positions and H, and a positive
// Gershgorin bound (1: ok; 0: a non-finite entry; -1: no positive bound).
int boundsUsable(const double* hb) {
  if (hb[kNonFinite] > 0.5) return 0;
  return (hb[6] > 0.0 && std::isfinite(hb[6]) && hb[9] > 0.0 && std::isfinite(hb[9])) ? 1 : -1;
}
void requireUsableBounds(const double* hb) {
  const int u = boundsUsable(hb);
  SG_REQUIRE(u != 0, SPHGPU_ERR_ARG, "non-finite node position or H");
  SG_REQUIRE(u > 0, SPHGPU_ERR_UNSUPPORTED, "neighbour search needs H with a positive Gershgorin bound on every node");
}

template <int D> __device__ __forceinline__ bool rowFinite(const double* x, const double* h) {
  bool ok = true;
  for (int k = 0; k < D; ++k) ok = ok && isfinite(x[k]);
  for (int k = 0; k < NSym<D>::value; ++k) ok = ok && isfinite(h[k]);
  return ok;
}

template <int D> __device__ double maxDiag(const double* h) {
  double m = h[0];
  for (int a = 1; a < D; ++a) m = dmax(m, h[sdiag<D>(a)]);
  return m;
}

template <int D>
__global__ void k_bounds(const double* __restrict__ pos, const double* __restrict__ H, int64_t n, double* partials) {
  __shared__ double sh[kNB][kBlock];
  double r[kNB] = SG_BOUNDS_IDENT;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    for (int k = 0; k < D; ++k) { const double x = pos[i * D + k]; r[k] = dmin(r[k], x); r[3 + k] = dmax(r[3 + k], x); }
    const double* h = H + i * NSym<D>::value;
    const double gm = gershgorinMin<D>(h);
    r[6] = dmin(r[6], gm);
    if (!isIsotropic<D>(h)) r[7] = 0.0;
    r[8] = dmax(r[8], maxDiag<D>(h));
    r[9] = dmax(r[9], gm);
    if (!rowFinite<D>(pos + i * D, h)) r[kNonFinite] = 1.0;
  }
  for (int k = 0; k < kNB; ++k) sh[k][threadIdx.x] = r[k];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < kNB; ++k) {
        const double a = sh[k][threadIdx.x], b = sh[k][threadIdx.x + s];
        sh[k][threadIdx.x] = isMaxSlot(k) ? dmax(a, b) : dmin(a, b);
      }
    __syncthreads();
  }
  if (threadIdx.x < kNB) partials[blockIdx.x * kNB + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_bounds_final(const double* partials, int nb, double* out) {
  __shared__ double sh[kNB][kBlock];
  double r[kNB] = SG_BOUNDS_IDENT;
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int k = 0; k < kNB; ++k) {
      const double v = partials[b * kNB + k];
      r[k] = isMaxSlot(k) ? dmax(r[k], v) : dmin(r[k], v);
    }
  for (int k = 0; k < kNB; ++k) sh[k][threadIdx.x] = r[k];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < kNB; ++k) {
        const double a = sh[k][threadIdx.x], b = sh[k][threadIdx.x + s];
        sh[k][threadIdx.x] = isMaxSlot(k) ? dmax(a, b) : dmin(a, b);
      }
    __syncthreads();
  }
  if (threadIdx.x < kNB) out[threadIdx.x] = sh[threadIdx.x][0];
}

struct DGrid {
  double lo[3];
  double cell;
  double invCell;
  int nc[3];
  int nb[3];
};

// Grid lists.  Grids [first[m], first[m+1]) serve ensemble member m (m = 0
// without an ensemble), in increasing cell size: the levels of a node set
// whose search radii span more than a factor 2.  A node lives in the first of
// its member's grids whose cell is at least its search radius bound
// ext/gershgorin(H)(1 + 1e-6).  g == null: one grid, passed by value.
struct SegGrids {
  const double* seg;        // member index per caller row, or null
  const DGrid* g;
  const int64_t* base;      // first global cell of each grid
  const int32_t* first;     // grids of each member, +1 entry
  double ext;
};

template <int D> __device__ __forceinline__ double searchRadius(const double* h, double ext) {
  return ext / gershgorinMin<D>(h) * (1.0 + 1.0e-6);
}

// Order-preserving map of a double onto uint64 (for atomic min/max).
__device__ __forceinline__ unsigned long long orderedBits(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
inline double fromOrderedBits(unsigned long long o) {
  const unsigned long long b = (o >> 63) ? (o & 0x7FFFFFFFFFFFFFFFull) : ~o;
  double x;
  std::memcpy(&x, &b, sizeof(x));
  return x;
}

__global__ void k_seg_bounds_init(unsigned long long* __restrict__ bits, int nSeg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nSeg * kNB) bits[i] = isMaxSlot(i % kNB) ? 0ull : ~0ull;
}

// The bounds of k_bounds per member.  A wavefront walks kSegGroups consecutive
// groups of 64 rows (members are contiguous runs of rows) and keeps a running
// reduction in registers while its groups stay in one member, with one atomic
// per value when the member changes or the walk ends; a group that spans
// members makes its atomics per lane.  Min and max are order-free, so the
// result does not depend on the grouping.
constexpr int kSegGroups = 16;
template <int D>
__global__ void k_seg_bounds(const double* __restrict__ pos, const double* __restrict__ H,
                             const double* __restrict__ seg, int64_t n, unsigned long long* __restrict__ bits) {
  const int lane = threadIdx.x & 63;
  const int64_t base = ((blockIdx.x * int64_t(blockDim.x) + threadIdx.x) >> 6) * int64_t(64 * kSegGroups);
  const double ident[kNB] = SG_BOUNDS_IDENT;
  double acc[kNB];
  for (int k = 0; k < kNB; ++k) acc[k] = ident[k];
  int mAcc = -1;                           // member of the running reduction (wave-uniform)
  auto atomicOne = [&](int m, int k, double v) {
    if (isMaxSlot(k)) atomicMax(bits + int64_t(m) * kNB + k, orderedBits(v));
    else atomicMin(bits + int64_t(m) * kNB + k, orderedBits(v));
  };
  auto flush = [&]() {                     // called by the whole wavefront
    for (int k = 0; k < kNB; ++k) {
      double v = acc[k];
      for (int sh = 32; sh > 0; sh >>= 1) {
        const double o = __shfl_xor(v, sh);
        v = isMaxSlot(k) ? dmax(v, o) : dmin(v, o);
      }
      if (lane == 0) atomicOne(mAcc, k, v);
      acc[k] = ident[k];
    }
  };
  for (int it = 0; it < kSegGroups; ++it) {
    const int64_t i = base + int64_t(it) * 64 + lane;
    double r[kNB];
    for (int k = 0; k < kNB; ++k) r[k] = ident[k];
    int m = -1;
    if (i < n) {
      m = int(seg[i]);
      for (int k = 0; k < D; ++k) { const double x = pos[i * D + k]; r[k] = x; r[3 + k] = x; }
      const double* h = H + i * NSym<D>::value;
      r[6] = gershgorinMin<D>(h);
      r[7] = isIsotropic<D>(h) ? 1.0 : 0.0;
      r[8] = maxDiag<D>(h);
      r[9] = r[6];
      r[kNonFinite] = rowFinite<D>(pos + i * D, h) ? 0.0 : 1.0;
    }
    const int m0 = __shfl(m, 0);
    if (m0 < 0) break;                     // lane 0 holds the group's first row: the rest are past n too
    if (__all(m == m0 || m < 0)) {
      if (m0 != mAcc) {
        if (mAcc >= 0) flush();
        mAcc = m0;
      }
      for (int k = 0; k < kNB; ++k) acc[k] = isMaxSlot(k) ? dmax(acc[k], r[k]) : dmin(acc[k], r[k]);
    } else {
      if (mAcc >= 0) flush();
      mAcc = -1;
      if (m >= 0)
        for (int k = 0; k < kNB; ++k) atomicOne(m, k, r[k]);
    }
  }
  if (mAcc >= 0) flush();
}

template <int D> __device__ __forceinline__ void cellCoords(const DGrid& g, const double* x, int* c) {
  c[0] = c[1] = c[2] = 0;
  for (int k = 0; k < D; ++k) {
    int ck = int((x[k] - g.lo[k]) * g.invCell);
    ck = ck < 0 ? 0 : (ck >= g.nc[k] ? g.nc[k] - 1 : ck);
    c[k] = ck;
  }
}

// Brick-ordered cell id: bricks of 4^nu cells in row-major order, Z-order inside.
template <int D> __device__ __forceinline__ int64_t cellId(const DGrid& g, int cx, int cy, int cz) {
  if constexpr (D == 1) {
    return cx;
  } else if constexpr (D == 2) {
    const int64_t b = int64_t(cy >> 2) * g.nb[0] + (cx >> 2);
    const int lx = cx & 3, ly = cy & 3;
    const int z = (lx & 1) | ((ly & 1) << 1) | ((lx & 2) << 1) | ((ly & 2) << 2);
    return b * 16 + z;
  } else {
    const int64_t b = (int64_t(cz >> 2) * g.nb[1] + (cy >> 2)) * g.nb[0] + (cx >> 2);
    const int lx = cx & 3, ly = cy & 3, lz = cz & 3;
    const int z = (lx & 1) | ((ly & 1) << 1) | ((lz & 1) << 2) | ((lx & 2) << 2) | ((ly & 2) << 3) | ((lz & 2) << 4);
    return b * 64 + z;
  }
}

template <int D>
__global__ void k_cell_count(const double* __restrict__ pos, const double* __restrict__ H, int64_t n, DGrid g,
                             SegGrids sg, int32_t* __restrict__ cellOf, int32_t* __restrict__ cellCount) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  int64_t base = 0;
  if (sg.g) {
    const int m = sg.seg ? int(sg.seg[i]) : 0;
    int q = sg.first[m];
    const int e = sg.first[m + 1];
    if (e - q > 1) {
      const double R = searchRadius<D>(H + i * NSym<D>::value, sg.ext);
      while (q < e - 1 && R > sg.g[q].cell) ++q;
    }
    g = sg.g[q];
    base = sg.base[q];
  }
  int c[3];
  cellCoords<D>(g, pos + i * D, c);
  const int32_t id = int32_t(base + cellId<D>(g, c[0], c[1], c[2]));
  cellOf[i] = id;
  atomicAdd(&cellCount[id], 1);
}

__global__ void k_cell_scatter(const int32_t* __restrict__ cellOf, const int32_t* __restrict__ cellStart,
                               int32_t* __restrict__ cellFill, int32_t* __restrict__ perm, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const int32_t c = cellOf[i];
  const int32_t slot = cellStart[c] + atomicAdd(&cellFill[c], 1);
  perm[slot] = int32_t(i);
}

// Order each cell's members by node index, or, with `xkey` (x of node i at
// xkey[i * stride]), by x and then node index, so that the members of a row
// of cells are in x order (insertion sort; cells hold a few to tens of nodes).
__global__ void k_cell_sort(const int32_t* __restrict__ cellStart, int32_t* __restrict__ perm, int64_t ncells,
                            const double* __restrict__ xkey, int stride) {
  const int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (c >= ncells) return;
  const int32_t b = cellStart[c], e = cellStart[c + 1];
  if (!xkey) {
    for (int32_t a = b + 1; a < e; ++a) {
      const int32_t v = perm[a];
      int32_t q = a - 1;
      while (q >= b && perm[q] > v) { perm[q + 1] = perm[q]; --q; }
      perm[q + 1] = v;
    }
    return;
  }
  constexpr int kSmall = 8;
  if (e - b <= kSmall) {
    // all keys loaded at once (independent loads), then an odd-even
    // transposition network in registers; padding sorts last
    int32_t id[kSmall];
    double xv[kSmall];
#pragma unroll
    for (int q = 0; q < kSmall; ++q) {
      id[q] = INT32_MAX;
      xv[q] = HUGE_VAL;
      if (b + q < e) {
        id[q] = perm[b + q];
        xv[q] = xkey[int64_t(id[q]) * stride];
      }
    }
#pragma unroll
    for (int r = 0; r < kSmall; ++r) {
#pragma unroll
      for (int q = r & 1; q + 1 < kSmall; q += 2) {
        const bool swap = xv[q] > xv[q + 1] || (xv[q] == xv[q + 1] && id[q] > id[q + 1]);
        const double x0 = xv[q], x1 = xv[q + 1];
        const int32_t i0 = id[q], i1 = id[q + 1];
        xv[q] = swap ? x1 : x0;
        xv[q + 1] = swap ? x0 : x1;
        id[q] = swap ? i1 : i0;
        id[q + 1] = swap ? i0 : i1;
      }
    }
#pragma unroll
    for (int q = 0; q < kSmall; ++q)
      if (b + q < e) perm[b + q] = id[q];
    return;
  }
  for (int32_t a = b + 1; a < e; ++a) {
    const int32_t v = perm[a];
    const double xv = xkey[int64_t(v) * stride];
    int32_t q = a - 1;
    for (; q >= b; --q) {
      const int32_t u = perm[q];
      const double xu = xkey[int64_t(u) * stride];
      if (xu < xv || (xu == xv && u < v)) break;
      perm[q + 1] = u;
    }
    perm[q + 1] = v;
  }
}

__global__ void k_iperm(const int32_t* __restrict__ perm, int32_t* __restrict__ iperm, int64_t n) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k < n) iperm[perm[k]] = int32_t(k);
}

// Row writer: keeps four entries in registers and stores them as one int4.
// Each entry is shifted in at w, so after every fourth the buffer holds them
// oldest first with no branch on the position; the last partial chunk is
// aligned once per row.  Entries past the row length are zero.  `at` walks
// the row's chunks (ellChunk(c, n, k) = c n + k).
struct RowWriter {
  int4 buf;
  int cnt;
  int cap;
  int64_t n;
  int4* at;
  __device__ __forceinline__ RowWriter(int4* o, int64_t n_, int64_t k_, int cap_)
      : buf{0, 0, 0, 0}, cnt(0), cap(cap_), n(n_), at(o + ellChunk(0, n_, k_)) {}
  __device__ __forceinline__ void push(int32_t v) {
    buf.x = buf.y;
    buf.y = buf.z;
    buf.z = buf.w;
    buf.w = v;
    ++cnt;
    if ((cnt & 3) == 0 && cnt <= cap) {
      *at = buf;
      at += n;
    }
  }
  // push(v) when acc, with the entry shifted in by selects: only the chunk
  // store is conditional
  // (ROOM: the caller has shown that the row stays within its capacity)
  template <bool ROOM = false>
  __device__ __forceinline__ void pushIf(bool acc, int32_t v) {
    buf.x = acc ? buf.y : buf.x;
    buf.y = acc ? buf.z : buf.y;
    buf.z = acc ? buf.w : buf.z;
    buf.w = acc ? v : buf.w;
    cnt += acc ? 1 : 0;
    if (acc && (cnt & 3) == 0 && (ROOM || cnt <= cap)) {
      *at = buf;
      at += n;
    }
  }
  __device__ __forceinline__ void finish() {
    const int r = cnt & 3;
    if (r == 0 || cnt > cap) return;
    const int4 o = r == 1 ? make_int4(buf.w, 0, 0, 0)
                          : (r == 2 ? make_int4(buf.z, buf.w, 0, 0) : make_int4(buf.y, buf.z, buf.w, 0));
    *at = o;
  }
};

// ---------------------------------------------------------------------------
// Staged brick search, 3-D, isotropic H.
// ---------------------------------------------------------------------------
constexpr int kStageEdge = 8;                       // 4-cell brick + 2 cells each side
constexpr int kStaged = kStageEdge * kStageEdge * kStageEdge;
constexpr int kCandCap = 2048;
constexpr int kRunRows = 9;      // one z slab: 5 runs, 3 sentinels, the read-ahead row

struct SearchArgs {
  DGrid g;
  double ext, ext2;
  float lo32, hi32;         // ext2*(1 -+ 1e-5): FP32 classification margins
  float prune2, cell32, invCell32;   // staged path: (search radius (1 + 1e-4))^2, cell edge and its inverse (FP32)
  int64_t n, nInt;
  int cap;
  int32_t* keepSlot;        // optional: set to 1 for every ghost slot some row holds
};

// Exact reference test for isotropic H: eta = h r componentwise (the SymTensor
// product with zero off-diagonals), magnitude2 left to right.
__device__ __forceinline__ bool exactIso(double rx, double ry, double rz, double hi, double hj, double ext2) {
  const double ax = hi * rx, ay = hi * ry, az = hi * rz;
  const double bx = hj * rx, by = hj * ry, bz = hj * rz;
  const double e2i = ax * ax + ay * ay + az * az;
  const double e2j = bx * bx + by * by + bz * bz;
  return (e2i <= ext2 || e2j <= ext2);
}

__global__ void __launch_bounds__(kBlock) k_search_staged(const double* __restrict__ rec, const int32_t* __restrict__ cellStart,
                                                         SearchArgs a, int32_t* __restrict__ rowCount,
                                                         int4* __restrict__ nbr, int32_t* __restrict__ flags) {
  using R = RL<true>;
  __shared__ int32_t sOff[kStaged + 1];
  __shared__ float4 sCand[kCandCap];               // x,y,z offsets from the brick origin, (1/h)^2 (FP32)
  __shared__ int32_t sSlot[kCandCap];
  // per lane: the runs of one z slab of its stencil (t0 | t1 << 16),
  // lane-strided; up to 5 runs, three sentinels and the row the reader reads
  // ahead of them.  Before the runs are formed the same memory holds each
  // staged cell's first slot and the scan's partial sums (52 KB in all: three
  // workgroups per CU)
  __shared__ uint32_t sRun[kRunRows * kBlock];
  static_assert(kRunRows * kBlock >= kStaged + kBlock, "sStart and sc alias the run table");
  int32_t* const sStart = reinterpret_cast<int32_t*>(sRun);
  int32_t* const sc = sStart + kStaged;
  const DGrid& g = a.g;
  const int b = int(xcdBlock(int64_t(g.nb[0]) * g.nb[1] * g.nb[2]));
  if (b >= g.nb[0] * g.nb[1] * g.nb[2]) return;   // uniform over the block
  const int bx = b % g.nb[0], by = (b / g.nb[0]) % g.nb[1], bz = b / (g.nb[0] * g.nb[1]);
  const int64_t h0 = cellStart[int64_t(b) * 64], h1 = cellStart[int64_t(b) * 64 + 64];
  if (h0 == h1) return;                           // uniform over the block
  const int cx0 = bx * 4 - 2, cy0 = by * 4 - 2, cz0 = bz * 4 - 2;
  const double ox = g.lo[0] + (bx * 4) * g.cell, oy = g.lo[1] + (by * 4) * g.cell, oz = g.lo[2] + (bz * 4) * g.cell;
  // staged cell table (x fastest)
  for (int t = threadIdx.x; t < kStaged; t += blockDim.x) {
    const int sx = t % kStageEdge, sy = (t / kStageEdge) % kStageEdge, sz = t / (kStageEdge * kStageEdge);
    const int cx = cx0 + sx, cy = cy0 + sy, cz = cz0 + sz;
    int32_t st = 0, cnt = 0;
    if (cx >= 0 && cx < g.nc[0] && cy >= 0 && cy < g.nc[1] && cz >= 0 && cz < g.nc[2]) {
      const int64_t id = cellId<3>(g, cx, cy, cz);
      st = cellStart[id];
      cnt = cellStart[id + 1] - st;
    }
    sStart[t] = st;
    sOff[t + 1] = cnt;
  }
  if (threadIdx.x == 0) sOff[0] = 0;
  __syncthreads();
  // inclusive scan of sOff[1..512] (two elements per thread)
  {
    const int t0 = 2 * threadIdx.x + 1, t1 = t0 + 1;
    const int32_t v0 = sOff[t0], v1 = sOff[t1];
    const int32_t pair = v0 + v1;
    sc[threadIdx.x] = pair;
    __syncthreads();
    for (int s = 1; s < kBlock; s <<= 1) {
      const int32_t add = threadIdx.x >= s ? sc[threadIdx.x - s] : 0;
      __syncthreads();
      sc[threadIdx.x] += add;
      __syncthreads();
    }
    const int32_t excl = sc[threadIdx.x] - pair;
    sOff[t0] = excl + v0;
    sOff[t1] = excl + v0 + v1;
  }
  __syncthreads();
  const int32_t total = sOff[kStaged];
  const bool staged = total <= kCandCap;
  int ghosts = 0;
  if (staged) {
    for (int t = threadIdx.x; t < kStaged; t += blockDim.x) {
      const int32_t o = sOff[t], cnt = sOff[t + 1] - o, st = sStart[t];
      for (int q = 0; q < cnt; ++q) {
        const double4 r = *reinterpret_cast<const double4*>(rec + int64_t(st + q) * R::STRIDE);
        sCand[o + q] = make_float4(float(r.x - ox), float(r.y - oy), float(r.z - oz), float(r.w) * float(r.w));
        // bit 31 flags a ghost (slots stay below 2^31)
        const bool ghost = rec[int64_t(st + q) * R::STRIDE + R::ORIG] >= double(a.nInt);
        ghosts |= ghost ? 1 : 0;
        sSlot[o + q] = int32_t(uint32_t(st + q) | (ghost ? 0x80000000u : 0u));
      }
    }
  } else if (threadIdx.x == 0) {
    atomicAdd(&flags[1], 1);                      // bricks that read candidates from global memory
  }
  __syncthreads();
  // ghost marks only where the staged candidates hold a ghost (the outer
  // bricks): a scalar flag, so interior bricks skip the mark code per candidate
  const bool marks =
      __builtin_amdgcn_readfirstlane(__syncthreads_or(ghosts)) != 0 && a.keepSlot != nullptr;

  // One lane per node; the node's 25 (dz, dy) runs of up to 5 staged cells
  // are contiguous candidate ranges in LDS.  Rows leave as int4 chunks, one
  // per lane, contiguous across a wavefront's slots.  Bricks whose candidates
  // exceed the stage read them from the records in global memory.
  // the cell table's first slots are read again only by bricks that are not
  // staged, which write no runs
  for (int64_t k = h0 + threadIdx.x; k < h1; k += blockDim.x) {
    const double* ri = rec + k * R::STRIDE;
    if (int64_t(ri[R::ORIG]) >= a.nInt) { rowCount[k] = 0; continue; }
    const double xi = ri[R::X], yi = ri[R::X + 1], zi = ri[R::X + 2], hi = ri[R::HINV];
    const float fx = float(xi - ox), fy = float(yi - oy), fz = float(zi - oz);
    const float hi2 = float(hi) * float(hi);
    int c[3];
    const double xv[3] = {xi, yi, zi};
    cellCoords<3>(g, xv, c);
    const int sx = c[0] - cx0, sy = c[1] - cy0, sz = c[2] - cz0;   // in [2, 5]
    RowWriter row(nbr, a.n, k, a.cap);
    if (staged) {
      for (int dz = -2; dz <= 2; ++dz) {
        // 1. the node's non-empty runs of this z slab, in dy order, into its
        //    column of the run table: the candidates of each (dz, dy) row of
        //    staged cells inside the x chord of a sphere of the search radius
        //    (with a 1e-4 margin) around the node
        int nr = 0, nCand = 0;
        {
          // distance from the node to the cell slab along z (staged cell t spans
          // [(t - 2) c, (t - 1) c) in brick-origin offsets)
          const float zlo = float(sz + dz - 2) * a.cell32, zhi = zlo + a.cell32;
          const float dzm = fz < zlo ? zlo - fz : (fz > zhi ? fz - zhi : 0.f);
          for (int dy = -2; dy <= 2; ++dy) {
            const float ylo = float(sy + dy - 2) * a.cell32, yhi = ylo + a.cell32;
            const float dym = fy < ylo ? ylo - fy : (fy > yhi ? fy - yhi : 0.f);
            const float rem2 = a.prune2 - (dzm * dzm + dym * dym);
            if (rem2 < 0.f) continue;
            const float rem = __builtin_sqrtf(rem2);
            int x0 = int(floorf((fx - rem) * a.invCell32)) + 2, x1 = int(floorf((fx + rem) * a.invCell32)) + 2;
            x0 = x0 < sx - 2 ? sx - 2 : x0;
            x1 = x1 > sx + 2 ? sx + 2 : x1;
            if (x0 > x1) continue;
            const int rowb = ((sz + dz) * kStageEdge + (sy + dy)) * kStageEdge;
            // the members of a cell are in x order: the chord [fx - rem, fx + rem]
            // starts and ends inside the end cells, found by two bisections
            int32_t t0 = sOff[rowb + x0], t1 = sOff[rowb + x1 + 1];
            {
              int32_t lo = t0, up = sOff[rowb + x0 + 1];
              const float key = fx - rem;
              while (lo < up) {
                const int32_t mid = (lo + up) >> 1;
                const bool below = sCand[mid].x < key;
                lo = below ? mid + 1 : lo;
                up = below ? up : mid;
              }
              t0 = lo;
            }
            {
              int32_t lo = sOff[rowb + x1], up = t1;
              const float key = fx + rem;
              while (lo < up) {
                const int32_t mid = (lo + up) >> 1;
                const bool inside = sCand[mid].x <= key;
                lo = inside ? mid + 1 : lo;
                up = inside ? up : mid;
              }
              t1 = lo;
            }
            if (t1 > t0) {
              sRun[nr++ * kBlock + threadIdx.x] = uint32_t(t0) | (uint32_t(t1) << 16);
              nCand += t1 - t0;
            }
          }
        }
        // 2. one walk over the slab's candidates: a wavefront iterates as often
        //    as its longest candidate list in the slab, not its longest run in
        //    every (dz, dy) row.  Candidates are read two ahead of their test (three
        //    register sets in rotation); the reader steps through the runs with
        //    no branch and, past the last candidate, through three sentinel runs
        //    (candidate 0), so every index it forms stays inside the tables.
        //    The rare paths (the exact FP64 test, ghost marks) are entered only
        //    when some lane of the wavefront needs them.
        // one instance of the walk with ghost marks and one without, chosen by
        // the block-uniform flag
        // plain: no lane of the wavefront can meet itself in this slab (only
        // the dz = 0 slab holds the node's own cell) or pass its row capacity
        // in it, so the walk skips both checks
        auto walk = [&](auto marksOn, auto plainOn) {
          constexpr bool plain = decltype(plainOn)::value;
          for (int q = 0; q < 3; ++q) sRun[(nr + q) * kBlock + threadIdx.x] = 0u;
          const uint32_t first = sRun[threadIdx.x];
          int32_t t = int32_t(first & 0xffffu), tEnd = int32_t(first >> 16);
          uint32_t nxt = sRun[kBlock + threadIdx.x];
          int r = 0;
          auto advance = [&]() {
            const bool same = t + 1 < tEnd;
            r = same ? r : r + 1;
            t = same ? t + 1 : int32_t(nxt & 0xffffu);
            tEnd = same ? tEnd : int32_t(nxt >> 16);
            nxt = sRun[(r + 1) * kBlock + threadIdx.x];
          };
          auto read = [&](float4& c, int32_t& raw) {
            c = sCand[t];
            raw = sSlot[t];
          };
          auto test = [&](const float4 cj, const int32_t raw) {
            const float rx = fx - cj.x, ry = fy - cj.y, rz = fz - cj.z;
            // FP32 estimate of min(h_i^2, h_j^2) |r|^2: fused products (any
            // rounding change stays inside the band that takes the exact test);
            // q min(a, b) = min(q a, q b) exactly for q >= 0
            const float q = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
            const float tmin = q * (cj.w < hi2 ? cj.w : hi2);
            bool in = tmin < a.lo32;
            const bool band = (tmin >= a.lo32) & (tmin <= a.hi32);
            const int32_t sj = raw & 0x7fffffff;
            // the exact test in the rare band (the branch is skipped when no
            // lane of the wavefront takes it)
            if (band) {
              const double* rj = rec + int64_t(sj) * R::STRIDE;
              in = exactIso(xi - rj[R::X], yi - rj[R::X + 1], zi - rj[R::X + 2], hi, rj[R::HINV], a.ext2);
            }
            const bool acc = in && (plain || sj != int32_t(k));
            row.template pushIf<plain>(acc, sj);
            if constexpr (decltype(marksOn)::value) {
              if (acc && raw < 0) a.keepSlot[sj] = 1;
            }
          };
          float4 c0, c1, c2;
          int32_t s0, s1, s2;
          read(c0, s0);
          advance();
          read(c1, s1);
          advance();
          for (int left = nCand;;) {
            read(c2, s2);
            advance();
            test(c0, s0);
            if (--left == 0) break;
            read(c0, s0);
            advance();
            test(c1, s1);
            if (--left == 0) break;
            read(c1, s1);
            advance();
            test(c2, s2);
            if (--left == 0) break;
          }
        };
        if (nr > 0) {
          const bool plain = dz != 0 && __all(row.cnt + nCand <= a.cap);
          if (marks) {
            if (plain) walk(std::true_type{}, std::true_type{});
            else walk(std::true_type{}, std::false_type{});
          } else {
            if (plain) walk(std::false_type{}, std::true_type{});
            else walk(std::false_type{}, std::false_type{});
          }
        }
      }
    } else {
      for (int dz = -2; dz <= 2; ++dz) {
        for (int dy = -2; dy <= 2; ++dy) {
          const int base = ((sz + dz) * kStageEdge + (sy + dy)) * kStageEdge + (sx - 2);
          for (int dx = 0; dx < 5; ++dx) {
            const int cidx = base + dx;
            const int32_t st = sStart[cidx], ncell = sOff[cidx + 1] - sOff[cidx];
            for (int32_t q2 = 0; q2 < ncell; ++q2) {
              const int32_t sj = st + q2;
              if (sj == int32_t(k)) continue;
              const double* rj = rec + int64_t(sj) * R::STRIDE;
              if (exactIso(xi - rj[R::X], yi - rj[R::X + 1], zi - rj[R::X + 2], hi, rj[R::HINV], a.ext2)) {
                row.push(sj);
                if (a.keepSlot && rj[R::ORIG] >= double(a.nInt)) a.keepSlot[sj] = 1;
              }
            }
          }
        }
      }
    }
    row.finish();
    rowCount[k] = row.cnt;
  }
}

// ---------------------------------------------------------------------------
// Global search: any dimension, any H (cell edge = R, 3^nu stencil).
// ---------------------------------------------------------------------------
template <int D, bool ISO>
__global__ void k_search_global(const double* __restrict__ rec, const int32_t* __restrict__ cellStart, SearchArgs a,
                                SegGrids sg, int32_t* __restrict__ rowCount, int4* __restrict__ nbr) {
  using R = RL<ISO>;
  const int64_t k = xcdBlock((a.n + kBlock - 1) / kBlock) * kBlock + threadIdx.x;
  if (k >= a.n) return;
  const double* ri_ = rec + k * R::STRIDE;
  const int64_t oi = int64_t(ri_[R::ORIG]);
  if (oi >= a.nInt) { rowCount[k] = 0; return; }
  int ga = 0, gb = 1;
  if (sg.g) {
    const int m = sg.seg ? int(sg.seg[oi]) : 0;
    ga = sg.first[m];
    gb = sg.first[m + 1];
  }
  constexpr int NS = NSym<D>::value;
  V<D> ri; S<D> Hi;
  for (int c = 0; c < D; ++c) ri.c[c] = ri_[R::X + c];
  if constexpr (ISO) {
    for (int c = 0; c < NS; ++c) Hi.c[c] = 0.0;
    for (int c = 0; c < D; ++c) Hi.c[sdiag<D>(c)] = ri_[RL<true>::HINV];
  } else {
    for (int c = 0; c < NS; ++c) Hi.c[c] = ri_[RL<false>::H + c];
  }
  // own search radius bound (only needed with several levels)
  double Ri = 0.0;
  if (gb - ga > 1) {
    if constexpr (ISO) Ri = a.ext / ri_[RL<true>::HINV] * (1.0 + 1.0e-6);
    else Ri = a.ext / gershgorinMin<D>(Hi.c) * (1.0 + 1.0e-6);
  }
  RowWriter row(nbr, a.n, k, a.cap);
  for (int q = ga; q < gb; ++q) {
  const DGrid g = sg.g ? sg.g[q] : a.g;
  const int64_t base = sg.g ? sg.base[q] : 0;
  // cells that can hold a neighbour: the 3^nu stencil when the grid's cell
  // covers both radii (nodes of a grid have radius <= its cell), else every
  // cell within Ri
  int c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0};
  {
    int cc[3];
    cellCoords<D>(g, ri.c, cc);
    for (int d = 0; d < D; ++d) {
      if (Ri <= g.cell) {
        c0[d] = cc[d] - 1 < 0 ? 0 : cc[d] - 1;
        c1[d] = cc[d] + 1 >= g.nc[d] ? g.nc[d] - 1 : cc[d] + 1;
      } else {
        const double t0 = floor((ri.c[d] - Ri - g.lo[d]) * g.invCell), t1 = floor((ri.c[d] + Ri - g.lo[d]) * g.invCell);
        c0[d] = t0 < 0.0 ? 0 : (t0 > double(g.nc[d] - 1) ? g.nc[d] - 1 : int(t0));
        c1[d] = t1 < 0.0 ? 0 : (t1 > double(g.nc[d] - 1) ? g.nc[d] - 1 : int(t1));
      }
    }
  }
  for (int z = c0[2]; z <= c1[2]; ++z) {
    for (int y = c0[1]; y <= c1[1]; ++y) {
      for (int x = c0[0]; x <= c1[0]; ++x) {
        const int64_t id = base + cellId<D>(g, x, y, z);
        const int32_t b = cellStart[id], e = cellStart[id + 1];
        for (int32_t j = b; j < e; ++j) {
          if (j == k) continue;
          const double* rj_ = rec + int64_t(j) * R::STRIDE;
          V<D> rij;
          for (int c = 0; c < D; ++c) rij.c[c] = ri.c[c] - rj_[R::X + c];
          S<D> Hj;
          if constexpr (ISO) {
            for (int c = 0; c < NS; ++c) Hj.c[c] = 0.0;
            for (int c = 0; c < D; ++c) Hj.c[sdiag<D>(c)] = rj_[RL<true>::HINV];
          } else {
            for (int c = 0; c < NS; ++c) Hj.c[c] = rj_[RL<false>::H + c];
          }
          const double eta2i = vmag2(sdot(Hi, rij));
          const double eta2j = vmag2(sdot(Hj, rij));
          if (eta2i <= a.ext2 || eta2j <= a.ext2) {
            row.push(j);
            if (a.keepSlot && rj_[R::ORIG] >= double(a.nInt)) a.keepSlot[j] = 1;
          }
        }
      }
    }
  }
  }
  row.finish();
  rowCount[k] = row.cnt;
}

// Internal slots in slot order: the work list of the density and pair passes.
__global__ void k_work_flag(const int32_t* __restrict__ perm, int64_t n, int64_t nInt, int32_t* __restrict__ flag) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k < n) flag[k] = (perm[k] >= 0 && perm[k] < nInt) ? 1 : 0;
}
__global__ void k_work_scatter(const int32_t* __restrict__ flag, const int32_t* __restrict__ scan, int64_t n,
                               int32_t* __restrict__ work) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k < n && flag[k]) work[scan[k]] = int32_t(k);
}

// Total entries and longest row.
__global__ void k_row_stats(const int32_t* __restrict__ rowCount, int64_t n, unsigned long long* __restrict__ total,
                            int32_t* __restrict__ maxRow) {
  __shared__ unsigned long long st[kBlock];
  __shared__ int32_t sm[kBlock];
  unsigned long long s = 0;
  int32_t m = 0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t c = rowCount[i];
    s += unsigned(c);
    m = c > m ? c : m;
  }
  st[threadIdx.x] = s;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      st[threadIdx.x] += st[threadIdx.x + w];
      sm[threadIdx.x] = sm[threadIdx.x] > sm[threadIdx.x + w] ? sm[threadIdx.x] : sm[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(total, st[0]);
    atomicMax(maxRow, sm[0]);
  }
}

// ---------------------------------------------------------------------------
// Reference row order: sort each row by caller index.  One wavefront per row,
// K keys per lane (rows up to 256 entries), bitonic network in registers;
// longer rows are sorted by a single lane.
// ---------------------------------------------------------------------------
template <int K, typename T>
__device__ __forceinline__ void waveBitonic(T (&key)[K], int lane) {
  constexpr int N = 64 * K;
#pragma unroll
  for (int kk = 2; kk <= N; kk <<= 1) {
#pragma unroll
    for (int jj = kk >> 1; jj > 0; jj >>= 1) {
      if (jj >= 64) {
        const int qj = jj / 64;
#pragma unroll
        for (int q = 0; q < K; ++q) {
          const int p = q ^ qj;
          if (p > q) {
            const int e = q * 64 + lane;
            const bool asc = (e & kk) == 0;
            const T a = key[q], b = key[p];
            const bool sw = asc ? (a > b) : (a < b);
            key[q] = sw ? b : a;
            key[p] = sw ? a : b;
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < K; ++q) {
          const int e = q * 64 + lane;
          const T other = __shfl_xor(key[q], jj);
          const bool asc = (e & kk) == 0;
          const bool lower = (lane & jj) == 0;
          const T mn = key[q] < other ? key[q] : other;
          const T mx = key[q] < other ? other : key[q];
          key[q] = (lower == asc) ? mn : mx;
        }
      }
    }
  }
}

template <int K>
__device__ void sortRowWave(int32_t* __restrict__ nbr, int64_t n, int64_t k, int len, const int32_t* __restrict__ perm,
                            const int32_t* __restrict__ iperm, int lane) {
  int32_t key[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const int e = q * 64 + lane;
    key[q] = e < len ? perm[nbr[ellIndex(e, n, k)]] : 0x7fffffff;
  }
  waveBitonic<K>(key, lane);
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const int e = q * 64 + lane;
    if (e < len) nbr[ellIndex(e, n, k)] = iperm[key[q]];
  }
}

// Several NodeLists: the order in which the reference's scatter delivers node
// x's pair terms (calculatePairInteraction, ConnectivityMapInline.hh:169-186):
// first the internal neighbours that own their pair (lower caller index, in
// that order), then x's own pairs in ConnectivityMap order, by (NodeList,
// index within the NodeList), a NodeList's ghosts after its internal nodes.
// With one NodeList this is ascending caller index.
// mapOrder: ConnectivityMap's own order, by (NodeList, index within the
// NodeList), for the loops that walk a node's full connectivity
// (iterateIdealH.cc:164-205) rather than receive from owners.
__device__ __forceinline__ int64_t receiveKey(int32_t c, int32_t x, const double* __restrict__ seg, int64_t nInt,
                                              bool mapOrder) {
  if (!mapOrder && c < nInt && c < x) return int64_t(c);
  const int64_t local = c < nInt ? int64_t(c) : (int64_t(1) << 31) + c;
  return (int64_t(1) << 62) | (int64_t(seg[c]) << 32) | local;
}
__device__ __forceinline__ int32_t callerOfKey(int64_t key) {
  if (key < (int64_t(1) << 62)) return int32_t(key);
  const int64_t low = key & 0xffffffffll;
  return int32_t(low >= (int64_t(1) << 31) ? low - (int64_t(1) << 31) : low);
}
template <int K>
__device__ void sortRowWaveLists(int32_t* __restrict__ nbr, int64_t n, int64_t k, int len,
                                 const int32_t* __restrict__ perm, const int32_t* __restrict__ iperm,
                                 const double* __restrict__ seg, int64_t nInt, bool mapOrder, int lane) {
  const int32_t x = perm[k];
  int64_t key[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const int e = q * 64 + lane;
    key[q] = e < len ? receiveKey(perm[nbr[ellIndex(e, n, k)]], x, seg, nInt, mapOrder) : INT64_MAX;
  }
  waveBitonic<K>(key, lane);
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const int e = q * 64 + lane;
    if (e < len) nbr[ellIndex(e, n, k)] = iperm[callerOfKey(key[q])];
  }
}

__global__ void __launch_bounds__(kBlock) k_row_sort(const int32_t* __restrict__ rowCount, int32_t* __restrict__ nbr,
                                                    int64_t n, const int32_t* __restrict__ perm,
                                                    const int32_t* __restrict__ iperm,
                                                    const double* __restrict__ seg, int64_t nInt, int mapOrder) {
  const int lane = threadIdx.x & 63;
  const int64_t k = (blockIdx.x * int64_t(blockDim.x) + threadIdx.x) >> 6;
  if (k >= n) return;                              // uniform per wavefront
  const int len = rowCount[k];
  if (len <= 1) return;
  if (seg) {
    const bool mo = mapOrder != 0;
    if (len <= 64) sortRowWaveLists<1>(nbr, n, k, len, perm, iperm, seg, nInt, mo, lane);
    else if (len <= 128) sortRowWaveLists<2>(nbr, n, k, len, perm, iperm, seg, nInt, mo, lane);
    else if (len <= 256) sortRowWaveLists<4>(nbr, n, k, len, perm, iperm, seg, nInt, mo, lane);
    else if (lane == 0) {
      const int32_t x = perm[k];
      for (int a2 = 1; a2 < len; ++a2) {
        const int32_t v = nbr[ellIndex(a2, n, k)];
        const int64_t pv = receiveKey(perm[v], x, seg, nInt, mo);
        int q = a2 - 1;
        while (q >= 0 && receiveKey(perm[nbr[ellIndex(q, n, k)]], x, seg, nInt, mo) > pv) {
          nbr[ellIndex(q + 1, n, k)] = nbr[ellIndex(q, n, k)];
          --q;
        }
        nbr[ellIndex(q + 1, n, k)] = v;
      }
    }
    return;
  }
  if (len <= 64) sortRowWave<1>(nbr, n, k, len, perm, iperm, lane);
  else if (len <= 128) sortRowWave<2>(nbr, n, k, len, perm, iperm, lane);
  else if (len <= 256) sortRowWave<4>(nbr, n, k, len, perm, iperm, lane);
  else if (lane == 0) {
    for (int a2 = 1; a2 < len; ++a2) {
      const int32_t v = nbr[ellIndex(a2, n, k)];
      const int32_t pv = perm[v];
      int q = a2 - 1;
      while (q >= 0 && perm[nbr[ellIndex(q, n, k)]] > pv) { nbr[ellIndex(q + 1, n, k)] = nbr[ellIndex(q, n, k)]; --q; }
      nbr[ellIndex(q + 1, n, k)] = v;
    }
  }
}

// ---- gathering a few int32 values for one readback -------------------------
struct Pick32 {
  int64_t at[16];
  int m;
};
__global__ void k_pick_i32(const int32_t* __restrict__ src, Pick32 p, int32_t* __restrict__ out) {
  const int q = threadIdx.x;
  if (q < p.m) out[q] = src[p.at[q]];
}

// ---- zeroing several int32 buffers in one dispatch ------------------------
__global__ void k_zero_many(ZeroSet z) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int q = 0; q < z.m; ++q)
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < z.n[q]; i += stride) z.p[q][i] = 0;
}

// ---- exclusive scans -------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

template <class TI, class TO>
__global__ void k_scan_tiles(const TI* __restrict__ in, TO* __restrict__ out, TO* __restrict__ tileSums, int64_t n) {
  __shared__ TO sh[kBlock];
  const int64_t base = blockIdx.x * int64_t(kScanTile) + threadIdx.x * kScanItems;
  TO v[kScanItems];
  TO local = 0;
  for (int q = 0; q < kScanItems; ++q) {
    const int64_t i = base + q;
    v[q] = (i < n) ? TO(in[i]) : TO(0);
    local += v[q];
  }
  sh[threadIdx.x] = local;
  __syncthreads();
  for (int s = 1; s < kBlock; s <<= 1) {
    TO t = (threadIdx.x >= s) ? sh[threadIdx.x - s] : TO(0);
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  TO run = sh[threadIdx.x] - local;
  for (int q = 0; q < kScanItems; ++q) {
    const int64_t i = base + q;
    if (i < n) out[i] = run;
    run += v[q];
  }
  if (threadIdx.x == kBlock - 1) tileSums[blockIdx.x] = sh[kBlock - 1];
}

template <class TO>
__global__ void k_scan_sums(TO* __restrict__ sums, int64_t m, TO* __restrict__ total) {
  __shared__ TO sh[kBlock];
  TO carry = 0;
  for (int64_t base = 0; base < m; base += kBlock) {
    const int64_t i = base + threadIdx.x;
    const TO v = (i < m) ? sums[i] : TO(0);
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 1; s < kBlock; s <<= 1) {
      TO t = (threadIdx.x >= s) ? sh[threadIdx.x - s] : TO(0);
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < m) sums[i] = carry + sh[threadIdx.x] - v;
    const TO blockTotal = sh[kBlock - 1];
    __syncthreads();
    carry += blockTotal;
  }
  if (threadIdx.x == 0) *total = carry;
}

template <class TO>
__global__ void k_scan_add(TO* __restrict__ out, const TO* __restrict__ sums, int64_t n, const TO* __restrict__ total) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) out[i] += sums[i / kScanTile];
  if (i == n) out[n] = *total;
}

// Small scans in one workgroup: tiles of kScanOneTile items in sequence with a
// running carry (one dispatch instead of three; small simulations run several
// scans per step and are bound by dispatch latency).
constexpr int kScanOneBlock = 1024;
constexpr int kScanOneTile = kScanOneBlock * kScanItems;
constexpr int64_t kScanOneMax = 4 * int64_t(kScanOneTile);

template <class TI, class TO>
__global__ void __launch_bounds__(kScanOneBlock) k_scan_one(const TI* __restrict__ in, TO* __restrict__ out,
                                                            int64_t n) {
  __shared__ TO sh[kScanOneBlock];
  TO carry = 0;
  for (int64_t tile = 0; tile < n; tile += kScanOneTile) {
    const int64_t base = tile + int64_t(threadIdx.x) * kScanItems;
    TO v[kScanItems];
    TO local = 0;
    for (int q = 0; q < kScanItems; ++q) {
      const int64_t i = base + q;
      v[q] = (i < n) ? TO(in[i]) : TO(0);
      local += v[q];
    }
    sh[threadIdx.x] = local;
    __syncthreads();
    for (int s = 1; s < kScanOneBlock; s <<= 1) {
      const TO t = (threadIdx.x >= s) ? sh[threadIdx.x - s] : TO(0);
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    TO run = carry + sh[threadIdx.x] - local;
    for (int q = 0; q < kScanItems; ++q) {
      const int64_t i = base + q;
      if (i < n) out[i] = run;
      run += v[q];
    }
    carry += sh[kScanOneBlock - 1];
    __syncthreads();                               // sh is rewritten by the next tile
  }
  if (threadIdx.x == 0) out[n] = carry;
}

template <class TI, class TO>
void exclusiveScan(const TI* in, TO* out, int64_t n, Sim& s) {
  // out has n+1 entries; out[n] = total.
  if (n <= kScanOneMax) {
    hipLaunchKernelGGL((k_scan_one<TI, TO>), dim3(1), dim3(kScanOneBlock), 0, s.stream, in, out, n);
    SG_CHECK_HIP(hipGetLastError());
    return;
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  s.scanTmp.ensure(sizeof(TO) * size_t(tiles + 2));
  TO* sums = s.scanTmp.as<TO>();
  TO* total = sums + tiles + 1;
  if (n > 0) hipLaunchKernelGGL((k_scan_tiles<TI, TO>), dim3(tiles), dim3(kBlock), 0, s.stream, in, out, sums, n);
  hipLaunchKernelGGL((k_scan_sums<TO>), dim3(1), dim3(kBlock), 0, s.stream, sums, tiles, total);
  hipLaunchKernelGGL((k_scan_add<TO>), dim3(nblocks(n + 1)), dim3(kBlock), 0, s.stream, out, sums, n, total);
  SG_CHECK_HIP(hipGetLastError());
}

// A grid of cell edge `cell` over the bounds hb: cells per axis, bricks, cell count.
template <int D>
void gridWithCell(const double* hb, double cell, Grid& G, DGrid& dg, int64_t& bricks) {
  G.cell = cell;
  dg = DGrid{};
  dg.cell = cell;
  dg.invCell = 1.0 / cell;
  bricks = 1;
  const int bw = (D == 1) ? 1 : 4;
  for (int k = 0; k < 3; ++k) {
    G.lo[k] = k < D ? hb[k] : 0.0;
    G.nc[k] = k < D ? int(std::floor((hb[3 + k] - hb[k]) / cell)) + 1 : 1;
    G.nb[k] = k < D ? (G.nc[k] + bw - 1) / bw : 1;
    dg.lo[k] = G.lo[k]; dg.nc[k] = G.nc[k]; dg.nb[k] = G.nb[k];
    if (k < D) bricks *= G.nb[k];
  }
  G.ncells = (D == 1) ? int64_t(G.nc[0]) : bricks * (D == 2 ? 16 : 64);
  SG_REQUIRE(G.ncells < (int64_t(1) << 31) - 2, SPHGPU_ERR_CAPACITY, "neighbour grid too large");
}

// Search grid from the bounds hb (k_bounds layout).  Staged path: cells of
// edge R/2 (5^3 stencil).  Global path: cells of edge >= R (3^nu stencil),
// enlarged if the grid would be too fine.
template <int D>
void gridFromBounds(const double* hb, double ext, bool allowStaged, Grid& G, DGrid& dg, int64_t& bricks,
                    bool& stagedPath) {
  requireUsableBounds(hb);
  const double R = ext / hb[6] * (1.0 + 1.0e-6);
  double spanX = 0.0;
  for (int k = 0; k < D; ++k) spanX = std::max(spanX, hb[3 + k] - hb[k]);
  const int maxCells = D == 3 ? 1024 : (D == 2 ? 32768 : (1 << 26));
  stagedPath = (D == 3) && allowStaged;
  double cell = 0.5 * R * (1.0 + 1.0e-9);
  if (stagedPath && spanX / cell > maxCells - 2) stagedPath = false;
  if (!stagedPath) {
    cell = R * (1.0 + 1.0e-9);
    if (spanX / cell > maxCells - 2) cell = spanX / (maxCells - 2);
  }
  G.span = stagedPath ? 2 : 1;
  G.R2 = R * R;
  gridWithCell<D>(hb, cell, G, dg, bricks);
}

// Levels of a node set whose search radii span more than a factor 2: cells
// of edge Rmin 2^L (>= R at the top level), each grid enlarged if too fine.
// One grid when the radii are within a factor 2.
template <int D>
void levelGrids(const double* hb, double ext, std::vector<DGrid>& out, std::vector<int64_t>& cells) {
  requireUsableBounds(hb);
  const double R = ext / hb[6] * (1.0 + 1.0e-6);
  const double Rmin = ext / hb[9] * (1.0 + 1.0e-6);
  Grid G;
  DGrid dg;
  int64_t bricks;
  if (!(R > 2.0 * Rmin)) {
    bool staged;
    gridFromBounds<D>(hb, ext, false, G, dg, bricks, staged);
    out.push_back(dg);
    cells.push_back(G.ncells);
    return;
  }
  double spanX = 0.0;
  for (int k = 0; k < D; ++k) spanX = std::max(spanX, hb[3 + k] - hb[k]);
  const int maxCells = D == 3 ? 1024 : (D == 2 ? 32768 : (1 << 26));
  double c = Rmin * (1.0 + 1.0e-9);
  for (int L = 0; L < 40; ++L) {
    const bool top = c >= R * (1.0 + 1.0e-9);
    double cell = c;
    if (spanX / cell > maxCells - 2) cell = spanX / (maxCells - 2);
    gridWithCell<D>(hb, cell, G, dg, bricks);
    out.push_back(dg);
    cells.push_back(G.ncells);
    if (top) break;
    c *= 2.0;
  }
}

// Grid lists to the device; returns the total cell count.
int64_t uploadGrids(Sim& s, const std::vector<DGrid>& grids, const std::vector<int64_t>& cells,
                    const std::vector<int32_t>& first) {
  hipStream_t st = s.stream;
  const size_t ng = grids.size();
  std::vector<int64_t> base(ng + 1, 0);
  for (size_t q = 0; q < ng; ++q) base[q + 1] = base[q] + cells[q];
  SG_REQUIRE(base.back() < (int64_t(1) << 31) - 2, SPHGPU_ERR_CAPACITY, "neighbour grid too large");
  // the grids, their first global cells and each member's first grid in one
  // buffer, staged in pinned memory at the same offsets and sent in one copy;
  // the next use of the staging waits for pinnedInDone, so no wait here
  const size_t bG = sizeof(DGrid) * ng, bB = sizeof(int64_t) * (ng + 1), bF = sizeof(int32_t) * first.size();
  const size_t oB = (bG + 63) / 64 * 64, oF = oB + (bB + 63) / 64 * 64;
  s.segGrid.ensure(oF + bF);
  char* d = static_cast<char*>(s.segGrid.p);
  s.gridList = d;
  s.gridCellBase = reinterpret_cast<int64_t*>(d + oB);
  s.gridFirst = reinterpret_cast<int32_t*>(d + oF);
  char* h = static_cast<char*>(hostStaging(s, 1, oF + bF));
  std::memcpy(h, grids.data(), bG);
  std::memcpy(h + oB, base.data(), bB);
  std::memcpy(h + oF, first.data(), bF);
  SG_CHECK_HIP(hipMemcpyAsync(d, h, oF + bF, hipMemcpyHostToDevice, st));
  if (!s.pinnedInDone) SG_CHECK_HIP(hipEventCreateWithFlags(&s.pinnedInDone, hipEventDisableTiming));
  SG_CHECK_HIP(hipEventRecord(s.pinnedInDone, st));
  return base.back();
}

template <int D>
void buildNeighborsD(Sim& s, double ext) {
  const int64_t n = s.nTot;
  hipStream_t st = s.stream;
  SG_CHECK_HIP(hipEventRecord(s.ev[0], st));

  // 1. bounds
  Grid& G = s.grid;
  DGrid dg{};
  int64_t bricks = 1;
  bool stagedPath = false;
  double hb[kNB];
  SegGrids sgr{nullptr, nullptr, nullptr, nullptr, ext};
  if (s.nSeg > 0) {
    const int ns = s.nSeg;
    s.segBoundsBits.ensure(sizeof(unsigned long long) * size_t(ns) * kNB);
    unsigned long long* bits = s.segBoundsBits.as<unsigned long long>();
    hipLaunchKernelGGL(k_seg_bounds_init, dim3(nblocks(int64_t(ns) * kNB)), dim3(kBlock), 0, st, bits, ns);
    hipLaunchKernelGGL((k_seg_bounds<D>), dim3(nblocks((n + kSegGroups - 1) / kSegGroups)), dim3(kBlock), 0, st,
                       s.pos.as<double>(), s.H.as<double>(), s.seg.as<double>(), n, bits);
    SG_CHECK_HIP(hipGetLastError());
    const size_t nbits = size_t(ns) * kNB;
    unsigned long long* hbits = static_cast<unsigned long long*>(hostStaging(s, 0, nbits * sizeof(unsigned long long)));
    SG_CHECK_HIP(hipMemcpyAsync(hbits, bits, nbits * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    SG_CHECK_HIP(hipStreamSynchronize(st));
    std::vector<DGrid> grids;
    std::vector<int64_t> cells;
    std::vector<int32_t> first(size_t(ns) + 1, 0);
    // a member with a non-finite position or H (or no positive bound of its
    // H) fails alone: the step is abandoned before anything of it is kept,
    // and the caller evicts the member and steps the others
    MemberFailure bad;
    for (int m = 0; m < ns; ++m) {
      double hm[kNB];
      for (int k = 0; k < kNB; ++k) hm[k] = fromOrderedBits(hbits[size_t(m) * kNB + k]);
      const int u = boundsUsable(hm);
      if (u <= 0) bad.add(m, u == 0 ? SPHGPU_MEMBER_NONFINITE : SPHGPU_MEMBER_BAD_H);
    }
    if (!bad.members.empty()) throw bad;
    bool iso = true;
    double absMax = 0.0;
    for (int m = 0; m < ns; ++m) {
      double hm[kNB];
      for (int k = 0; k < kNB; ++k) hm[k] = fromOrderedBits(hbits[size_t(m) * kNB + k]);
      for (int k = 0; k < D; ++k) absMax = std::max(absMax, std::max(std::fabs(hm[k]), std::fabs(hm[3 + k])));
      iso = iso && hm[7] > 0.5;
      levelGrids<D>(hm, ext, grids, cells);
      first[size_t(m) + 1] = int32_t(grids.size());
      if (m == 0) for (int k = 0; k < kNB; ++k) hb[k] = hm[k];
      hb[6] = std::min(hb[6], hm[6]);
      hb[8] = std::max(hb[8], hm[8]);
    }
    s.isotropic = iso && !s.hasMultipliers;
    s.rMax = ext / hb[6];
    s.posAbsMax = absMax;
    G = Grid{};
    G.ncells = uploadGrids(s, grids, cells, first);
    sgr = SegGrids{s.seg.as<double>(), static_cast<const DGrid*>(s.gridList), s.gridCellBase, s.gridFirst, ext};
  } else {
    const int nbB = int(std::min<int64_t>(1024, nblocks(n)));
    s.reduceTmp.ensure(sizeof(double) * (size_t(nbB) * kNB + 16));
    double* partials = s.reduceTmp.as<double>();
    double* bounds = partials + size_t(nbB) * kNB;
    hipLaunchKernelGGL((k_bounds<D>), dim3(nbB), dim3(kBlock), 0, st, s.pos.as<double>(), s.H.as<double>(), n, partials);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kBlock), 0, st, partials, nbB, bounds);
    SG_CHECK_HIP(hipGetLastError());
    double* hbp = static_cast<double*>(hostStaging(s, 0, sizeof(hb)));
    SG_CHECK_HIP(hipMemcpyAsync(hbp, bounds, sizeof(hb), hipMemcpyDeviceToHost, st));
    SG_CHECK_HIP(hipStreamSynchronize(st));
    std::memcpy(hb, hbp, sizeof(hb));
    requireUsableBounds(hb);
    s.isotropic = hb[7] > 0.5 && !s.hasMultipliers;
    s.rMax = ext / hb[6];
    s.posAbsMax = 0.0;
    for (int k = 0; k < D; ++k) s.posAbsMax = std::max(s.posAbsMax, std::max(std::fabs(hb[k]), std::fabs(hb[3 + k])));
    std::vector<DGrid> grids;
    std::vector<int64_t> cells;
    levelGrids<D>(hb, ext, grids, cells);
    if (grids.size() == 1) {
      gridFromBounds<D>(hb, ext, s.isotropic, G, dg, bricks, stagedPath);
    } else {
      // smoothing scales spanning more than a factor 2: levels
      std::vector<int32_t> first = {0, int32_t(grids.size())};
      G = Grid{};
      G.ncells = uploadGrids(s, grids, cells, first);
      sgr = SegGrids{nullptr, static_cast<const DGrid*>(s.gridList), s.gridCellBase, s.gridFirst, ext};
    }
  }

  // 2. counting sort by cell, then by node index (staged search: by x) inside a cell
  s.cellOf.ensure(sizeof(int32_t) * size_t(n));
  s.cellCount.ensure(sizeof(int32_t) * size_t(G.ncells + 1));
  s.cellFill.ensure(sizeof(int32_t) * size_t(G.ncells + 1));
  s.cellStart.ensure(sizeof(int32_t) * size_t(G.ncells + 2));
  s.perm.ensure(sizeof(int32_t) * size_t(n));
  s.iperm.ensure(sizeof(int32_t) * size_t(n));
  // zero the cell counters, the search flags and the ghost-use marks in one dispatch
  const bool markUse = s.markGhostUse && s.nTot > s.nInt;
  if (markUse) s.slotKeep.ensure(sizeof(int32_t) * size_t(n + 1));
  s.flags.ensure(sizeof(int32_t) * 8 + sizeof(unsigned long long) * 2);
  {
    ZeroSet z{};
    z.add(s.cellCount.as<int32_t>(), G.ncells + 1);
    z.add(s.cellFill.as<int32_t>(), G.ncells + 1);
    z.add(s.flags.as<int32_t>(), int64_t(s.flags.bytes / sizeof(int32_t)));
    if (markUse) z.add(s.slotKeep.as<int32_t>(), n);
    zeroMany(z, s.stream);
  }
  hipLaunchKernelGGL((k_cell_count<D>), dim3(nblocks(n)), dim3(kBlock), 0, st, s.pos.as<double>(), s.H.as<double>(),
                     n, dg, sgr,
                     s.cellOf.as<int32_t>(), s.cellCount.as<int32_t>());
  exclusiveScan<int32_t, int32_t>(s.cellCount.as<int32_t>(), s.cellStart.as<int32_t>(), G.ncells, s);
  hipLaunchKernelGGL(k_cell_scatter, dim3(nblocks(n)), dim3(kBlock), 0, st, s.cellOf.as<int32_t>(),
                     s.cellStart.as<int32_t>(), s.cellFill.as<int32_t>(), s.perm.as<int32_t>(), n);
  // the staged search cuts each row of cells to the chord of its sphere by x
  hipLaunchKernelGGL(k_cell_sort, dim3(nblocks(G.ncells)), dim3(kBlock), 0, st, s.cellStart.as<int32_t>(),
                     s.perm.as<int32_t>(), G.ncells, stagedPath ? s.pos.as<double>() : nullptr, D);
  hipLaunchKernelGGL(k_iperm, dim3(nblocks(n)), dim3(kBlock), 0, st, s.perm.as<int32_t>(), s.iperm.as<int32_t>(), n);
  SG_CHECK_HIP(hipGetLastError());
  s.nSlots = n;
  // the records the next kernel reads, written with the FP64 record: the
  // caller's hint, else the density sum's
  packRecords(s, s.isotropic ? (s.buildRecHint >= 0 ? s.buildRecHint : densityRecordSec(s)) : 0);

  // 3. rows
  SearchArgs sa{};
  sa.g = dg;
  sa.ext = ext;
  sa.ext2 = ext * ext;
  // FP32 classification band.  Offsets from the brick origin are at most 3R, so
  // |r| carries an absolute error below 7.2e-7 R; near the threshold of the
  // smallest smoothing scale (|r| = ext/hmax_inv >= R*hinv_min/hinv_max) that is
  // a relative error of 1.44e-6*ratio in |r|^2, plus ~5e-7 from the FP32
  // products.  The band is four times that, and never narrower than 1e-5.
  const double ratio = hb[8] / hb[6];
  const double band = std::max(1.0e-5, 4.0 * (1.44e-6 * ratio + 5.0e-7));
  sa.lo32 = float(sa.ext2 * (1.0 - band));
  sa.hi32 = float(sa.ext2 * (1.0 + band));
  {
    const double Rm = ext / hb[6] * (1.0 + 1.0e-6) * (1.0 + 1.0e-4);
    sa.prune2 = float(Rm * Rm);
    sa.cell32 = float(dg.cell);
    sa.invCell32 = float(1.0 / dg.cell);
  }
  sa.n = n;
  sa.nInt = s.nInt;
  sa.keepSlot = nullptr;
  if (markUse) sa.keepSlot = s.slotKeep.as<int32_t>();
  if (s.ellCap <= 0) s.ellCap = D == 3 ? 160 : (D == 2 ? 48 : 16);
  s.rowCount.ensure(sizeof(int32_t) * size_t(n));
  unsigned long long* dTotal = reinterpret_cast<unsigned long long*>(s.flags.as<int32_t>() + 4);
  for (int attempt = 0; attempt < 2; ++attempt) {
    sa.cap = s.ellCap;
    s.nbr.ensure(ellBytes(n, s.ellCap));
    if (attempt > 0) {
      // the first attempt's flags were zeroed with the cell counters; ghost
      // marks of the first attempt are a subset of the second's
      ZeroSet z{};
      z.add(s.flags.as<int32_t>(), int64_t(s.flags.bytes / sizeof(int32_t)));
      zeroMany(z, s.stream);
    }
    if (stagedPath) {
      hipLaunchKernelGGL(k_search_staged, dim3(int(xcdGrid(bricks))), dim3(kBlock), 0, st, s.rec.as<double>(),
                         s.cellStart.as<int32_t>(), sa, s.rowCount.as<int32_t>(), s.nbr.as<int4>(),
                         s.flags.as<int32_t>());
    } else if (s.isotropic) {
      hipLaunchKernelGGL((k_search_global<D, true>), dim3(int(xcdGrid(nblocks(n)))), dim3(kBlock), 0, st, s.rec.as<double>(),
                         s.cellStart.as<int32_t>(), sa, sgr, s.rowCount.as<int32_t>(), s.nbr.as<int4>());
    } else {
      hipLaunchKernelGGL((k_search_global<D, false>), dim3(int(xcdGrid(nblocks(n)))), dim3(kBlock), 0, st, s.rec.as<double>(),
                         s.cellStart.as<int32_t>(), sa, sgr, s.rowCount.as<int32_t>(), s.nbr.as<int4>());
    }
    hipLaunchKernelGGL(k_row_stats, dim3(int(std::min<int64_t>(1024, nblocks(n)))), dim3(kBlock), 0, st,
                       s.rowCount.as<int32_t>(), n, dTotal, s.flags.as<int32_t>());
    SG_CHECK_HIP(hipGetLastError());
    struct Flags { int32_t f[4]; unsigned long long total; };
    char* hp = static_cast<char*>(hostStaging(s, 0, 256 + sizeof(int32_t) * 16));
    // the ghost cull's first half rides on the same wait
    const bool cullHere = markUse && s.subStart.size() < 16;
    if (cullHere) cullLaunch(s, reinterpret_cast<int32_t*>(hp + 256));
    Flags* hfp = reinterpret_cast<Flags*>(hp);
    SG_CHECK_HIP(hipMemcpyAsync(hfp, s.flags.p, sizeof(Flags), hipMemcpyDeviceToHost, st));
    SG_CHECK_HIP(hipStreamSynchronize(st));
    const Flags hf = *hfp;
    if (cullHere) {
      const int32_t* cs = reinterpret_cast<const int32_t*>(hp + 256);
      s.cullStarts.assign(cs, cs + s.subStart.size() + 1);
      s.cullStartsReady = true;
    }
    s.maxRow = hf.f[0];
    s.stagedFallbacks = hf.f[1];
    s.nEntries = int64_t(hf.total);
    if (s.maxRow <= s.ellCap) break;
    SG_REQUIRE(attempt == 0, SPHGPU_ERR_CAPACITY, "neighbour row capacity exceeded twice");
    s.ellCap = ((s.maxRow + 15) / 16) * 16 + 16;
  }
  // 4. reference row order
  if (s.rowOrder == SPHGPU_ROW_ORDER_REFERENCE && n > 0)
    hipLaunchKernelGGL(k_row_sort, dim3(nblocks(n * 64)), dim3(kBlock), 0, st, s.rowCount.as<int32_t>(),
                       s.nbr.as<int32_t>(), n, s.perm.as<int32_t>(), s.iperm.as<int32_t>(),
                       s.nLists > 1 ? s.seg.as<double>() : nullptr, s.nInt, s.rowsMapOrder ? 1 : 0);
  // 5. work list of internal slots
  s.workFlag.ensure(sizeof(int32_t) * size_t(n + 1));
  s.workScan.ensure(sizeof(int32_t) * size_t(n + 1));
  s.work.ensure(sizeof(int32_t) * size_t(std::max<int64_t>(s.nInt, 1)));
  hipLaunchKernelGGL(k_work_flag, dim3(nblocks(n)), dim3(kBlock), 0, st, s.perm.as<int32_t>(), n, s.nInt,
                     s.workFlag.as<int32_t>());
  exclusiveScan<int32_t, int32_t>(s.workFlag.as<int32_t>(), s.workScan.as<int32_t>(), n, s);
  hipLaunchKernelGGL(k_work_scatter, dim3(nblocks(n)), dim3(kBlock), 0, st, s.workFlag.as<int32_t>(),
                     s.workScan.as<int32_t>(), n, s.work.as<int32_t>());
  SG_CHECK_HIP(hipGetLastError());
  SG_CHECK_HIP(hipEventRecord(s.ev[1], st));
}

}  // namespace

void exclusiveScanI32toI64(const int32_t* in, int64_t* out, int64_t n, Sim& s) { exclusiveScan<int32_t, int64_t>(in, out, n, s); }
void exclusiveScanI32(const int32_t* in, int32_t* out, int64_t n, Sim& s) { exclusiveScan<int32_t, int32_t>(in, out, n, s); }
void* hostStaging(Sim& s, int which, size_t bytes) {
  void*& p = which ? s.pinnedIn : s.pinnedOut;
  size_t& cap = which ? s.pinnedInBytes : s.pinnedOutBytes;
  // uploads: the copies of the previous use must have read the buffer
  if (which && s.pinnedInDone) SG_CHECK_HIP(hipEventSynchronize(s.pinnedInDone));
  if (bytes > cap) {
    SG_CHECK_HIP(hipStreamSynchronize(s.stream));   // no copy still uses the old buffer
    if (p) SG_CHECK_HIP(hipHostFree(p));
    p = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(4096, ((bytes + 4095) / 4096) * 4096);
    SG_CHECK_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
    cap = want;
  }
  return p;
}

void pickInts(Sim& s, const int32_t* src, const int64_t* at, int m, int32_t* hostOut) {
  SG_REQUIRE(m >= 1 && m <= 16, SPHGPU_ERR_ARG, "pickInts takes 1 to 16 values");
  Pick32 pk{};
  for (int q = 0; q < m; ++q) pk.at[q] = at[q];
  pk.m = m;
  s.pick.ensure(sizeof(int32_t) * 16);
  hipLaunchKernelGGL(k_pick_i32, dim3(1), dim3(64), 0, s.stream, src, pk, s.pick.as<int32_t>());
  SG_CHECK_HIP(hipGetLastError());
  SG_CHECK_HIP(hipMemcpyAsync(hostOut, s.pick.p, sizeof(int32_t) * size_t(m), hipMemcpyDeviceToHost, s.stream));
}

const int32_t* readInts(Sim& s, const int32_t* src, const int64_t* at, int m) {
  int32_t* h = static_cast<int32_t*>(hostStaging(s, 0, sizeof(int32_t) * 16));
  pickInts(s, src, at, m, h);
  SG_CHECK_HIP(hipStreamSynchronize(s.stream));
  return h;
}

void zeroMany(const ZeroSet& z, hipStream_t st) {
  int64_t most = 0;
  for (int q = 0; q < z.m; ++q) most = std::max(most, z.n[q]);
  if (most == 0) return;
  hipLaunchKernelGGL(k_zero_many, dim3(std::min(nblocks(most), 1024)), dim3(kBlock), 0, st, z);
  SG_CHECK_HIP(hipGetLastError());
}

void buildNeighbors(Sim& s, double ext) {
  s.cullStartsReady = false;
  SG_REQUIRE(s.nTot > 0, SPHGPU_ERR_STATE, "no nodes");
  SG_REQUIRE(ext > 0.0, SPHGPU_ERR_ARG, "kernel extent must be positive");
  s.kernelExtent = ext;
  s.neighborsValid = false;
  s.pairsValid = false;
  s.derivsValid = false;
  if (s.dim == 1) buildNeighborsD<1>(s, ext);
  else if (s.dim == 2) buildNeighborsD<2>(s, ext);
  else buildNeighborsD<3>(s, ext);
  s.timingPending |= 1u;
  s.neighborsValid = true;
}

}  // namespace sphgpu
