// Device ConnectivityMap build.
//
// Contract (src/Neighbor/ConnectivityMap.cc:630-749): for an internal node i, j
// is a neighbour iff |H_i r_ij|^2 <= ext^2 or |H_j r_ij|^2 <= ext^2 (FP64,
// r_ij = r_i - r_j, SymTensor.Vector and magnitude2 summed left to right), j != i,
// and each row is sorted ascending by node index (:735-738).  How candidates are
// found is free (the reference's NestedGridNeighbor only has to return a
// superset), so this uses one uniform grid:
//
//  1. bounds: min/max position and the smallest Gershgorin lower bound of
//     lambda_min(H) over all nodes (internal + ghost).  |H r| <= ext implies
//     |r| <= ext/lambda_min(H), so a cell edge R = ext/min(lambda) (plus a 2e-6
//     relative margin) makes the 3^nu cells around a node a superset.
//  2. counting sort of nodes by cell id.  Cells are numbered brick by brick
//     (4^nu cells per brick, Z-order inside a brick) so that a wavefront's nodes
//     and their neighbours are close in memory.  Each cell is then ordered by
//     node index, so the sorted order depends only on positions and indices.
//  3. per sorted node: count pass, scan, fill pass over the 3^nu cells with a
//     |r|^2 <= R^2 prefilter followed by the exact test.
//  4. per row: bitonic sort in LDS by node index (rows store sorted slots).
#include "internal.hh"
#include "dgeom.hh"

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

// partials[b*7 + {0..2 lo, 3..5 hi, 6 min lambda}]
template <int D>
__global__ void k_bounds(const double* __restrict__ pos, const double* __restrict__ H, int64_t n, double* partials) {
  __shared__ double sh[7][kBlock];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, lm = 1e300;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    for (int k = 0; k < D; ++k) { const double x = pos[i * D + k]; lo[k] = dmin(lo[k], x); hi[k] = dmax(hi[k], x); }
    lm = dmin(lm, gershgorinMin<D>(H + i * NSym<D>::value));
  }
  for (int k = 0; k < 3; ++k) { sh[k][threadIdx.x] = lo[k]; sh[3 + k][threadIdx.x] = hi[k]; }
  sh[6][threadIdx.x] = lm;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      for (int k = 0; k < 3; ++k) {
        sh[k][threadIdx.x] = dmin(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
        sh[3 + k][threadIdx.x] = dmax(sh[3 + k][threadIdx.x], sh[3 + k][threadIdx.x + s]);
      }
      sh[6][threadIdx.x] = dmin(sh[6][threadIdx.x], sh[6][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 7) partials[blockIdx.x * 7 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_bounds_final(const double* partials, int nb, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double r[7] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300, 1e300};
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 7; ++k) {
      const double v = partials[b * 7 + k];
      r[k] = (k >= 3 && k < 6) ? dmax(r[k], v) : dmin(r[k], v);
    }
  for (int k = 0; k < 7; ++k) out[k] = r[k];
}

struct DGrid {
  double lo[3];
  double invCell;
  double R2;
  int nc[3];
  int nb[3];
};

template <int D> __device__ __forceinline__ void cellCoords(const DGrid& g, const double* x, int* c) {
  c[0] = c[1] = c[2] = 0;
  for (int k = 0; k < D; ++k) {
    int ck = int((x[k] - g.lo[k]) * g.invCell);
    ck = ck < 0 ? 0 : (ck >= g.nc[k] ? g.nc[k] - 1 : ck);
    c[k] = ck;
  }
}

// Brick-ordered cell id: bricks of 4^nu cells in row-major order, Z-order inside.
template <int D> __device__ __forceinline__ int64_t cellId(const DGrid& g, const int* c) {
  if constexpr (D == 1) {
    return c[0];
  } else if constexpr (D == 2) {
    const int64_t b = int64_t(c[1] >> 2) * g.nb[0] + (c[0] >> 2);
    const int lx = c[0] & 3, ly = c[1] & 3;
    const int z = (lx & 1) | ((ly & 1) << 1) | ((lx & 2) << 1) | ((ly & 2) << 2);
    return b * 16 + z;
  } else {
    const int64_t b = (int64_t(c[2] >> 2) * g.nb[1] + (c[1] >> 2)) * g.nb[0] + (c[0] >> 2);
    const int lx = c[0] & 3, ly = c[1] & 3, lz = c[2] & 3;
    const int z = (lx & 1) | ((ly & 1) << 1) | ((lz & 1) << 2) | ((lx & 2) << 2) | ((ly & 2) << 3) | ((lz & 2) << 4);
    return b * 64 + z;
  }
}

template <int D>
__global__ void k_cell_count(const double* __restrict__ pos, int64_t n, DGrid g, int32_t* __restrict__ cellOf,
                             int32_t* __restrict__ cellCount) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  int c[3];
  cellCoords<D>(g, pos + i * D, c);
  const int32_t id = int32_t(cellId<D>(g, c));
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

// Order each cell's members by node index (insertion sort; cells hold tens of nodes).
__global__ void k_cell_sort(const int32_t* __restrict__ cellStart, int32_t* __restrict__ perm, int64_t ncells) {
  const int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (c >= ncells) return;
  const int32_t b = cellStart[c], e = cellStart[c + 1];
  for (int32_t a = b + 1; a < e; ++a) {
    const int32_t v = perm[a];
    int32_t q = a - 1;
    while (q >= b && perm[q] > v) { perm[q + 1] = perm[q]; --q; }
    perm[q + 1] = v;
  }
}

template <int D>
__global__ void k_gather_xh(const int32_t* __restrict__ perm, const double* __restrict__ pos, const double* __restrict__ H,
                            double* __restrict__ xs, double* __restrict__ Hs, int32_t* __restrict__ iperm, int64_t n) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k >= n) return;
  const int32_t i = perm[k];
  for (int c = 0; c < D; ++c) xs[k * D + c] = pos[int64_t(i) * D + c];
  constexpr int NS = NSym<D>::value;
  for (int c = 0; c < NS; ++c) Hs[k * NS + c] = H[int64_t(i) * NS + c];
  iperm[i] = int32_t(k);
}

// Exact ConnectivityMap test for the pair (i, j) in i's frame.
template <int D> __device__ __forceinline__ bool exactTest(const V<D>& rij, const S<D>& Hi, const S<D>& Hj, double ext2) {
  const double eta2i = vmag2(sdot(Hi, rij));
  const double eta2j = vmag2(sdot(Hj, rij));
  return (eta2i <= ext2 || eta2j <= ext2);
}

template <int D, bool FILL>
__global__ void k_rows(const double* __restrict__ xs, const double* __restrict__ Hs, const int32_t* __restrict__ perm,
                       const int32_t* __restrict__ cellStart, DGrid g, double ext2, int64_t n, int64_t nInt,
                       int32_t* __restrict__ rowCount, const int64_t* __restrict__ offsets, int32_t* __restrict__ nbr) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k >= n) return;
  if (perm[k] >= nInt) { if (!FILL) rowCount[k] = 0; return; }
  constexpr int NS = NSym<D>::value;
  V<D> ri; S<D> Hi;
  for (int c = 0; c < D; ++c) ri.c[c] = xs[k * D + c];
  for (int c = 0; c < NS; ++c) Hi.c[c] = Hs[k * NS + c];
  int cc[3];
  cellCoords<D>(g, ri.c, cc);
  int32_t count = 0;
  int64_t w = FILL ? offsets[k] : 0;
  const int64_t wEnd = FILL ? offsets[k + 1] : 0;   // the count pass sized this row
  const int dz0 = D >= 3 ? -1 : 0, dz1 = D >= 3 ? 1 : 0;
  const int dy0 = D >= 2 ? -1 : 0, dy1 = D >= 2 ? 1 : 0;
  for (int dz = dz0; dz <= dz1; ++dz) {
    const int z = cc[2] + dz;
    if (z < 0 || z >= g.nc[2]) continue;
    for (int dy = dy0; dy <= dy1; ++dy) {
      const int y = cc[1] + dy;
      if (y < 0 || y >= g.nc[1]) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int x = cc[0] + dx;
        if (x < 0 || x >= g.nc[0]) continue;
        const int c3[3] = {x, y, z};
        const int64_t id = cellId<D>(g, c3);
        const int32_t b = cellStart[id], e = cellStart[id + 1];
        for (int32_t j = b; j < e; ++j) {
          if (j == k) continue;
          V<D> rij;
          for (int c = 0; c < D; ++c) rij.c[c] = ri.c[c] - xs[int64_t(j) * D + c];
          if (vmag2(rij) > g.R2) continue;
          S<D> Hj;
          for (int c = 0; c < NS; ++c) Hj.c[c] = Hs[int64_t(j) * NS + c];
          if (exactTest<D>(rij, Hi, Hj, ext2)) {
            if (FILL && w < wEnd) nbr[w++] = j;
            ++count;
          }
        }
      }
    }
  }
  if (!FILL) rowCount[k] = count;
}

// Sort each row (sorted slots) ascending by node index.  One 64-lane block per
// row; rows up to kRowCap entries are sorted in LDS, longer ones by lane 0.
constexpr int kRowCap = 1024;
__global__ void __launch_bounds__(64) k_row_sort(const int64_t* __restrict__ offsets, const int32_t* __restrict__ perm,
                                                 int32_t* __restrict__ nbr, int64_t n) {
  __shared__ uint64_t key[kRowCap];
  const int64_t row = blockIdx.x;
  if (row >= n) return;
  const int64_t b = offsets[row], e = offsets[row + 1];
  const int len = int(e - b);
  if (len <= 1) return;
  if (len > kRowCap) {
    if (threadIdx.x == 0) {
      for (int64_t a = b + 1; a < e; ++a) {
        const int32_t v = nbr[a];
        const int32_t pv = perm[v];
        int64_t q = a - 1;
        while (q >= b && perm[nbr[q]] > pv) { nbr[q + 1] = nbr[q]; --q; }
        nbr[q + 1] = v;
      }
    }
    return;
  }
  int np2 = 1;
  while (np2 < len) np2 <<= 1;
  for (int t = threadIdx.x; t < np2; t += 64) {
    if (t < len) {
      const uint32_t slot = uint32_t(nbr[b + t]);
      key[t] = (uint64_t(uint32_t(perm[slot])) << 32) | slot;
    } else {
      key[t] = ~uint64_t(0);
    }
  }
  __syncthreads();
  for (int kk = 2; kk <= np2; kk <<= 1) {
    for (int jj = kk >> 1; jj > 0; jj >>= 1) {
      for (int t = threadIdx.x; t < np2; t += 64) {
        const int p = t ^ jj;
        if (p > t) {
          const uint64_t a = key[t], c = key[p];
          const bool up = (t & kk) == 0;
          if ((a > c) == up) { key[t] = c; key[p] = a; }
        }
      }
      __syncthreads();
    }
  }
  for (int t = threadIdx.x; t < len; t += 64) nbr[b + t] = int32_t(uint32_t(key[t] & 0xffffffffu));
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
  TO run = sh[threadIdx.x] - local;   // exclusive prefix of this thread
  for (int q = 0; q < kScanItems; ++q) {
    const int64_t i = base + q;
    if (i < n) out[i] = run;
    run += v[q];
  }
  if (threadIdx.x == kBlock - 1) tileSums[blockIdx.x] = sh[kBlock - 1];
}

// Single-block exclusive scan of the tile sums (in place); writes the total.
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

template <class TI, class TO>
void exclusiveScan(const TI* in, TO* out, int64_t n, Sim& s) {
  // out has n+1 entries; out[n] = total.
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  s.scanTmp.ensure(sizeof(TO) * size_t(tiles + 2));
  TO* sums = s.scanTmp.as<TO>();
  TO* total = sums + tiles + 1;
  if (n > 0) hipLaunchKernelGGL((k_scan_tiles<TI, TO>), dim3(tiles), dim3(kBlock), 0, s.stream, in, out, sums, n);
  hipLaunchKernelGGL((k_scan_sums<TO>), dim3(1), dim3(kBlock), 0, s.stream, sums, tiles, total);
  hipLaunchKernelGGL((k_scan_add<TO>), dim3(nblocks(n + 1)), dim3(kBlock), 0, s.stream, out, sums, n, total);
  SG_CHECK_HIP(hipGetLastError());
}

template <int D>
void buildNeighborsD(Sim& s, double ext) {
  const int64_t n = s.nTot;
  constexpr int NS = NSym<D>::value;
  hipStream_t st = s.stream;
  SG_CHECK_HIP(hipEventRecord(s.ev[0], st));

  // 1. bounds
  const int nbB = int(std::min<int64_t>(1024, nblocks(n)));
  s.reduceTmp.ensure(sizeof(double) * (size_t(nbB) * 7 + 8));
  double* partials = s.reduceTmp.as<double>();
  double* bounds = partials + size_t(nbB) * 7;
  hipLaunchKernelGGL((k_bounds<D>), dim3(nbB), dim3(kBlock), 0, st, s.pos.as<double>(), s.H.as<double>(), n, partials);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(64), 0, st, partials, nbB, bounds);
  SG_CHECK_HIP(hipGetLastError());
  double hb[7];
  SG_CHECK_HIP(hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, st));
  SG_CHECK_HIP(hipStreamSynchronize(st));
  SG_REQUIRE(hb[6] > 0.0 && std::isfinite(hb[6]), SPHGPU_ERR_UNSUPPORTED,
             "neighbour search needs H with a positive Gershgorin bound on every node");
  for (int k = 0; k < D; ++k)
    SG_REQUIRE(std::isfinite(hb[k]) && std::isfinite(hb[3 + k]), SPHGPU_ERR_ARG, "non-finite node position");

  Grid& G = s.grid;
  const double R = ext / hb[6] * (1.0 + 1.0e-6);
  double cell = R * (1.0 + 1.0e-9);
  double span = 0.0;
  for (int k = 0; k < D; ++k) span = std::max(span, hb[3 + k] - hb[k]);
  const int maxCells = D == 3 ? 1024 : (D == 2 ? 32768 : (1 << 26));
  if (span / cell > maxCells - 2) cell = span / (maxCells - 2);
  G.cell = cell;
  G.R2 = R * R;
  DGrid dg{};
  dg.invCell = 1.0 / cell;
  dg.R2 = G.R2;
  int64_t bricks = 1;
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

  // 2. counting sort by cell, then by node index inside a cell
  s.cellOf.ensure(sizeof(int32_t) * size_t(n));
  s.cellCount.ensure(sizeof(int32_t) * size_t(G.ncells + 1));
  s.cellFill.ensure(sizeof(int32_t) * size_t(G.ncells + 1));
  s.cellStart.ensure(sizeof(int32_t) * size_t(G.ncells + 2));
  s.perm.ensure(sizeof(int32_t) * size_t(n));
  s.iperm.ensure(sizeof(int32_t) * size_t(n));
  s.xs.ensure(sizeof(double) * size_t(n) * D);
  s.Hs.ensure(sizeof(double) * size_t(n) * NS);
  SG_CHECK_HIP(hipMemsetAsync(s.cellCount.p, 0, sizeof(int32_t) * size_t(G.ncells + 1), st));
  SG_CHECK_HIP(hipMemsetAsync(s.cellFill.p, 0, sizeof(int32_t) * size_t(G.ncells + 1), st));
  hipLaunchKernelGGL((k_cell_count<D>), dim3(nblocks(n)), dim3(kBlock), 0, st, s.pos.as<double>(), n, dg,
                     s.cellOf.as<int32_t>(), s.cellCount.as<int32_t>());
  exclusiveScan<int32_t, int32_t>(s.cellCount.as<int32_t>(), s.cellStart.as<int32_t>(), G.ncells, s);
  hipLaunchKernelGGL(k_cell_scatter, dim3(nblocks(n)), dim3(kBlock), 0, st, s.cellOf.as<int32_t>(),
                     s.cellStart.as<int32_t>(), s.cellFill.as<int32_t>(), s.perm.as<int32_t>(), n);
  hipLaunchKernelGGL(k_cell_sort, dim3(nblocks(G.ncells)), dim3(kBlock), 0, st, s.cellStart.as<int32_t>(),
                     s.perm.as<int32_t>(), G.ncells);
  hipLaunchKernelGGL((k_gather_xh<D>), dim3(nblocks(n)), dim3(kBlock), 0, st, s.perm.as<int32_t>(),
                     s.pos.as<double>(), s.H.as<double>(), s.xs.as<double>(), s.Hs.as<double>(),
                     s.iperm.as<int32_t>(), n);
  SG_CHECK_HIP(hipGetLastError());

  // 3. rows: count, scan, fill
  const double ext2 = ext * ext;
  s.rowCount.ensure(sizeof(int32_t) * size_t(n));
  s.offsets.ensure(sizeof(int64_t) * size_t(n + 1));
  hipLaunchKernelGGL((k_rows<D, false>), dim3(nblocks(n)), dim3(kBlock), 0, st, s.xs.as<double>(), s.Hs.as<double>(),
                     s.perm.as<int32_t>(), s.cellStart.as<int32_t>(), dg, ext2, n, s.nInt,
                     s.rowCount.as<int32_t>(), (const int64_t*)nullptr, (int32_t*)nullptr);
  exclusiveScan<int32_t, int64_t>(s.rowCount.as<int32_t>(), s.offsets.as<int64_t>(), n, s);
  int64_t total = 0;
  SG_CHECK_HIP(hipMemcpyAsync(&total, s.offsets.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SG_CHECK_HIP(hipStreamSynchronize(st));
  s.nEntries = total;
  s.nbr.ensure(sizeof(int32_t) * size_t(std::max<int64_t>(total, 1)));
  hipLaunchKernelGGL((k_rows<D, true>), dim3(nblocks(n)), dim3(kBlock), 0, st, s.xs.as<double>(), s.Hs.as<double>(),
                     s.perm.as<int32_t>(), s.cellStart.as<int32_t>(), dg, ext2, n, s.nInt,
                     s.rowCount.as<int32_t>(), s.offsets.as<int64_t>(), s.nbr.as<int32_t>());
  // 4. order rows by node index
  if (n > 0)
    hipLaunchKernelGGL(k_row_sort, dim3(n), dim3(64), 0, st, s.offsets.as<int64_t>(), s.perm.as<int32_t>(),
                       s.nbr.as<int32_t>(), n);
  SG_CHECK_HIP(hipGetLastError());
  SG_CHECK_HIP(hipEventRecord(s.ev[1], st));
}

}  // namespace

void exclusiveScanI32toI64(const int32_t* in, int64_t* out, int64_t n, Sim& s) { exclusiveScan<int32_t, int64_t>(in, out, n, s); }
void exclusiveScanI32(const int32_t* in, int32_t* out, int64_t n, Sim& s) { exclusiveScan<int32_t, int32_t>(in, out, n, s); }

void regatherSorted(Sim& s) {
  const int64_t n = s.nTot;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nblocks(n)), dim3(kBlock), 0, s.stream, s.perm.as<int32_t>(), s.pos.as<double>(),
                       s.H.as<double>(), s.xs.as<double>(), s.Hs.as<double>(), s.iperm.as<int32_t>(), n);
  };
  if (s.dim == 1) go(k_gather_xh<1>);
  else if (s.dim == 2) go(k_gather_xh<2>);
  else go(k_gather_xh<3>);
  SG_CHECK_HIP(hipGetLastError());
  s.xsStale = false;
}

void buildNeighbors(Sim& s, double ext) {
  SG_REQUIRE(s.nTot > 0, SPHGPU_ERR_STATE, "no nodes");
  SG_REQUIRE(ext > 0.0, SPHGPU_ERR_ARG, "kernel extent must be positive");
  s.kernelExtent = ext;
  s.neighborsValid = false;
  s.derivsValid = false;
  if (s.dim == 1) buildNeighborsD<1>(s, ext);
  else if (s.dim == 2) buildNeighborsD<2>(s, ext);
  else buildNeighborsD<3>(s, ext);
  float ms = 0.f;
  SG_CHECK_HIP(hipEventSynchronize(s.ev[1]));
  SG_CHECK_HIP(hipEventElapsedTime(&ms, s.ev[0], s.ev[1]));
  s.timing.neighbors_ms = ms;
  s.neighborsValid = true;
  s.xsStale = false;
}

}  // namespace sphgpu
