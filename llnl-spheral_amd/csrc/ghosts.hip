// Planar ghost boundaries on the device.
//
// The reference (src/Integrator/Integrator.cc:443-595) rebuilds its ghosts once
// per step: every boundary, in order, looks at all nodes that exist when it
// runs (internal nodes and the ghosts of the boundaries before it), selects
// its control nodes and appends one ghost per control
// (src/Boundary/PlanarBoundary.cc:112-182, 329-400):
//   hmax     = max over nodes i with min(|d_exit(x_i)|, |d_enter(x_i)|) < ext*h_i
//              of h_i = 1/lambda_min(H_i)
//   controls = nodes with 0 <= d_exit(x_i)/hmax <= ext, in index order
//   ghost    = mapPositionThroughPlanes(x_c, exit, enter); H, v by R = I - 2nn
//              (reflecting) or copied (periodic); scalars copied.
// After the connectivity is built, ghosts that no internal node sees and that
// control no other ghost are culled, the rest keep their order
// (src/Integrator/Integrator.cc:491-570, src/Boundary/Boundary.cc:182-233).
//
// Device form: per sub-boundary one pass over the current nodes computes hmax
// (wavefront max + one 64-bit atomic max per wavefront on the bit pattern of a
// non-negative double), a second flags the controls, an exclusive scan places
// the ghosts, and a third writes them after the current rows.  Ghost fields
// are filled boundary by boundary (a ghost's control always precedes it).
// Culling: the row search marks every ghost slot it puts in a row; a scan of
// the keep flags in caller order gives the old -> new index map; the state is
// compacted and the slot permutation renumbered.  Slots themselves stay where
// they are (culled ghosts become dead slots that no row holds), so the rows,
// their entries and their order are untouched and no second build is needed.
//
// Floating-point order follows the reference exactly: plane distances
// n.(x - p), closest points x + ((p - x).n) n (3-D), the 2-D and 1-D special
// cases, and the closed-form eigenvalues of SymTensor::eigenValues
// (src/Geometry/GeomSymmetricTensorInline_default.hh:2053-2110).  The 3-D
// eigenvalues call atan2, cos and sin; their device and host results may
// differ in the last bit, which moves hmax by far less than one ulp except in
// rare rounding ties.
#include "internal.hh"
#include "dgeom.hh"
#include "records.hh"
#include "ell.hh"
#include "eigen.hh"
#include <cmath>
#include <algorithm>

namespace sphgpu {

void exclusiveScanI32(const int32_t* in, int32_t* out, int64_t n, Sim& s);

namespace {

constexpr int kBlock = 256;
inline int nblocks(int64_t n, int bs = kBlock) { return int(std::max<int64_t>(1, (n + bs - 1) / bs)); }

struct DSub {
  double enterP[3], enterN[3], exitP[3], exitN[3];
  int reflect;
  double R[9];
};

DSub toDevice(const SubBoundary& b) {
  DSub d;
  for (int k = 0; k < 3; ++k) {
    d.enterP[k] = b.enterP[k]; d.enterN[k] = b.enterN[k]; d.exitP[k] = b.exitP[k]; d.exitN[k] = b.exitN[k];
  }
  d.reflect = b.reflect;
  for (int k = 0; k < 9; ++k) d.R[k] = b.R[k];
  return d;
}

template <int D> __device__ __forceinline__ V<D> loadV(const double* __restrict__ a, int64_t i) {
  V<D> v;
  for (int k = 0; k < D; ++k) v.c[k] = a[i * D + k];
  return v;
}
template <int D> __device__ __forceinline__ V<D> arr(const double* a) {
  V<D> v;
  for (int k = 0; k < D; ++k) v.c[k] = a[k];
  return v;
}
// GeomPlane::signedDistance: normal.dot(point - mPoint)
template <int D> __device__ __forceinline__ double signedDist(const double* P, const double* N, const V<D>& x) {
  return vdot(arr<D>(N), vsub(x, arr<D>(P)));
}
// GeomPlane::closestPointOnPlane (GeomPlane.cc:237-259, pointDistances.hh:53-57)
template <int D> __device__ __forceinline__ V<D> closestPoint(const double* P, const double* N, const V<D>& x) {
  if constexpr (D == 1) {
    return arr<1>(P);
  } else if constexpr (D == 2) {
    const V<2> a0p = vsub(x, arr<2>(P));
    V<2> dir;
    dir.c[0] = -N[1]; dir.c[1] = N[0];
    const double s = vdot(a0p, dir);
    V<2> o;
    for (int k = 0; k < 2; ++k) o.c[k] = P[k] + dir.c[k] * s;
    return o;
  } else {
    const double s = vdot(vsub(arr<3>(P), x), arr<3>(N));
    V<3> o;
    for (int k = 0; k < 3; ++k) o.c[k] = x.c[k] + N[k] * s;
    return o;
  }
}
// PlanarBoundary::updateGhostNodes: mapPositionThroughPlanes(x, exitPlane, enterPlane)
//   = closestPoint(enter, x) - signedDistance(exit, x) * enter.normal
template <int D> __device__ __forceinline__ V<D> ghostPosition(const DSub& b, const V<D>& x) {
  const V<D> p0 = closestPoint<D>(b.enterP, b.enterN, x);
  const double s = signedDist<D>(b.exitP, b.exitN, x);
  V<D> o;
  for (int k = 0; k < D; ++k) o.c[k] = p0.c[k] - b.enterN[k] * s;
  return o;
}

// R.v, rows left to right (GeomTensorInline_default.hh:1918-1922)
template <int D> __device__ __forceinline__ V<D> reflectV(const double* R, const V<D>& v) {
  V<D> o;
  for (int i = 0; i < D; ++i) {
    double s = R[i * D] * v.c[0];
    for (int k = 1; k < D; ++k) s = s + R[i * D + k] * v.c[k];
    o.c[i] = s;
  }
  return o;
}
// (R*(H*R)).Symmetric() (ReflectingBoundary.cc:155-175)
template <int D> __device__ __forceinline__ void reflectH(const double* R, const double* h, double* out) {
  double full[D * D], t1[D * D], t2[D * D];
  int c = 0;
  for (int i = 0; i < D; ++i)
    for (int j = i; j < D; ++j, ++c) { full[i * D + j] = h[c]; full[j * D + i] = h[c]; }
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      double s = full[i * D] * R[j];
      for (int k = 1; k < D; ++k) s = s + full[i * D + k] * R[k * D + j];
      t1[i * D + j] = s;
    }
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      double s = R[i * D] * t1[j];
      for (int k = 1; k < D; ++k) s = s + R[i * D + k] * t1[k * D + j];
      t2[i * D + j] = s;
    }
  c = 0;
  for (int i = 0; i < D; ++i)
    for (int j = i; j < D; ++j, ++c) out[c] = (i == j) ? t2[i * D + i] : 0.5 * (t2[i * D + j] + t2[j * D + i]);
}

// 1/lambda_min(H) of the internal nodes (a ghost's is computed when it is made).
template <int D>
__global__ void k_hinv_min(const double* __restrict__ H, int64_t n, double* __restrict__ hmaxOf) {
  constexpr int NS = NSym<D>::value;
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) hmaxOf[i] = 1.0 / minEigen<D>(H + i * NS);
}

// hmax of the nodes near either plane, one value: grid-stride loop, a block
// maximum and one 64-bit atomic max per block.
// Row counts of the ghost kernels: `n` rows, or, with nDev, the count the
// device holds (n is then the capacity the grid covers).
__device__ __forceinline__ int64_t liveRows(int64_t n, const int64_t* nDev) { return nDev ? *nDev : n; }

template <int D>
__global__ void k_ghost_hmax_one(const double* __restrict__ pos, const double* __restrict__ hmaxOf, int64_t n, DSub b,
                                 double ext, unsigned long long* __restrict__ hmaxBits, const int64_t* __restrict__ nDev) {
  __shared__ double wmax[kBlock / 64];
  double v = 0.0;
  const int64_t nn = liveRows(n, nDev);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < nn; i += int64_t(gridDim.x) * blockDim.x) {
    const V<D> x = loadV<D>(pos, i);
    const double hmaxi = hmaxOf[i];
    const double d = dmin(dabs(signedDist<D>(b.exitP, b.exitN, x)), dabs(signedDist<D>(b.enterP, b.enterN, x)));
    if (hmaxi > 0.0 && d < ext * hmaxi) v = dmax(v, hmaxi);
  }
  for (int sh = 32; sh > 0; sh >>= 1) v = dmax(v, __shfl_xor(v, sh));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) v = dmax(v, wmax[w]);
    if (v > 0.0) atomicMax(hmaxBits, (unsigned long long)__double_as_longlong(v));
  }
}

// hmax of the nodes near either plane per ensemble member (hmaxBits[member]).
template <int D>
__global__ void k_ghost_hmax(const double* __restrict__ pos, const double* __restrict__ hmaxOf,
                             const double* __restrict__ seg, int64_t n, DSub b, double ext,
                             unsigned long long* __restrict__ hmaxBits, const int64_t* __restrict__ nDev) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  double v = 0.0;
  int m = 0;
  if (i < liveRows(n, nDev)) {
    const V<D> x = loadV<D>(pos, i);
    const double hmaxi = hmaxOf[i];
    const double d = dmin(dabs(signedDist<D>(b.exitP, b.exitN, x)), dabs(signedDist<D>(b.enterP, b.enterN, x)));
    if (hmaxi > 0.0 && d < ext * hmaxi) v = hmaxi;
    if (seg) m = int(seg[i]);
  }
  const int m0 = __shfl(m, 0);
  if (__all(m == m0)) {
    for (int sh = 32; sh > 0; sh >>= 1) v = dmax(v, __shfl_xor(v, sh));
    if ((threadIdx.x & 63) == 0 && v > 0.0) atomicMax(hmaxBits + m0, (unsigned long long)__double_as_longlong(v));
  } else if (v > 0.0) {
    atomicMax(hmaxBits + m, (unsigned long long)__double_as_longlong(v));
  }
}

template <int D>
// blockCount (optional): each workgroup's count of flags, for k_ghost_make<BLK>.
__global__ void k_ghost_flag(const double* __restrict__ pos, const double* __restrict__ seg, int64_t n, DSub b,
                             double ext, const unsigned long long* __restrict__ hmaxBits, int32_t* __restrict__ flags,
                             const int64_t* __restrict__ nDev, int32_t* __restrict__ blockCount) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  int32_t f = 0;
  if (i < liveRows(n, nDev)) {                    // capacity rows beyond the live count take no flag
    const double hmax = __longlong_as_double((long long)hmaxBits[seg ? int(seg[i]) : 0]);
    if (hmax > 0.0) {
      const double disti = signedDist<D>(b.exitP, b.exitN, loadV<D>(pos, i)) / hmax;
      f = (disti >= 0.0 && disti <= ext) ? 1 : 0;
    }
  }
  if (i < n) flags[i] = f;
  if (blockCount) {
    const int c = __syncthreads_count(f);        // every thread of the block reaches it
    if (threadIdx.x == 0) blockCount[blockIdx.x] = c;
  }
}

// BLK: offs holds the exclusive scan of the flag kernel's per-workgroup counts
// (same grid and workgroup size), and a row's offset is its workgroup's plus
// its rank among the workgroup's flags: the global exclusive scan of flags.
// Otherwise offs is that scan itself.
template <int D, bool BLK>
__global__ void __launch_bounds__(kBlock) k_ghost_make(double* __restrict__ pos, double* __restrict__ H, double* __restrict__ seg, int64_t n,
                             int64_t nInt, const int32_t* __restrict__ flags, const int32_t* __restrict__ offs, DSub b,
                             int subIdx, int32_t* __restrict__ control, int32_t* __restrict__ subOf,
                             double* __restrict__ hmaxOf, const int64_t* __restrict__ nDev, int64_t cap,
                             int64_t* __restrict__ overflow) {
  constexpr int NS = NSym<D>::value;
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int64_t nn = liveRows(n, nDev);
  int64_t off;
  if constexpr (BLK) {
    const bool f = i < nn && flags[i] != 0;
    const unsigned long long m = __ballot(f);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ int waveCount[kBlock / 64];
    if (lane == 0) waveCount[wave] = __popcll(m);
    __syncthreads();
    if (!f) return;
    int before = 0;
    for (int w = 0; w < wave; ++w) before += waveCount[w];
    off = int64_t(offs[blockIdx.x]) + before + __popcll(m & ((1ull << lane) - 1ull));
  } else {
    if (i >= nn || !flags[i]) return;
    off = offs[i];
  }
  const int64_t g = nn + off;
  if (nDev && g >= cap) { *overflow = 1; return; }   // the capacity guess was short: the build redoes it exactly
  const V<D> xg = ghostPosition<D>(b, loadV<D>(pos, i));
  for (int k = 0; k < D; ++k) pos[g * D + k] = xg.c[k];
  if (b.reflect) {
    double hg[NS];
    reflectH<D>(b.R, H + i * NS, hg);
    for (int k = 0; k < NS; ++k) H[g * NS + k] = hg[k];
    hmaxOf[g] = 1.0 / minEigen<D>(hg);
  } else {
    for (int k = 0; k < NS; ++k) H[g * NS + k] = H[i * NS + k];
    hmaxOf[g] = hmaxOf[i];
  }
  control[g - nInt] = int32_t(i);
  subOf[g - nInt] = subIdx;
  if (seg) seg[g] = seg[i];
}

// After a plane with device-held counts: the plane's first ghost, then the new
// live count (the count of flags is the scan's total over the capacity).
struct GenDev {
  int64_t n;            // live rows
  int64_t overflow;     // set when a ghost fell beyond the capacity
  int64_t start[16];    // first ghost (offset from nInt) of each sub-boundary
};
// Both also clear the hmax words of the next plane.
__global__ void k_gen_init(GenDev* g, int64_t nInt, unsigned long long* __restrict__ hbits, int nh) {
  for (int q = threadIdx.x; q < nh; q += blockDim.x) hbits[q] = 0ull;
  if (threadIdx.x == 0) {
    g->n = nInt;
    g->overflow = 0;
  }
}
__global__ void k_gen_advance(GenDev* g, int b, const int32_t* __restrict__ total, int64_t nInt, int64_t cap,
                              unsigned long long* __restrict__ hbits, int nh) {
  for (int q = threadIdx.x; q < nh; q += blockDim.x) hbits[q] = 0ull;
  if (threadIdx.x == 0) {
    g->start[b] = g->n - nInt;
    const int64_t next = g->n + int64_t(*total);
    if (next > cap) g->overflow = 1;
    g->n = next > cap ? cap : next;
  }
}

struct FieldPtrs {
  double *pos, *H, *vel, *mass, *rho, *eps, *P, *cs, *omega, *fCl, *fCq, *fsh;
};

// Ghost rows [g0, g1) of one sub-boundary from their controls.
template <int D>
__global__ void k_ghost_fields(FieldPtrs f, int64_t g0, int64_t g1, int64_t nInt, const int32_t* __restrict__ control,
                               DSub b, int geometry) {
  constexpr int NS = NSym<D>::value;
  const int64_t g = g0 + blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (g >= g1) return;
  const int64_t c = control[g - nInt];
  if (geometry) {
    const V<D> xg = ghostPosition<D>(b, loadV<D>(f.pos, c));
    for (int k = 0; k < D; ++k) f.pos[g * D + k] = xg.c[k];
    if (b.reflect) {
      double hg[NS];
      reflectH<D>(b.R, f.H + c * NS, hg);
      for (int k = 0; k < NS; ++k) f.H[g * NS + k] = hg[k];
    } else {
      for (int k = 0; k < NS; ++k) f.H[g * NS + k] = f.H[c * NS + k];
    }
  }
  const V<D> v = loadV<D>(f.vel, c);
  const V<D> vg = b.reflect ? reflectV<D>(b.R, v) : v;
  for (int k = 0; k < D; ++k) f.vel[g * D + k] = vg.c[k];
  f.mass[g] = f.mass[c];
  f.rho[g] = f.rho[c];
  f.eps[g] = f.eps[c];
  f.P[g] = f.P[c];
  f.cs[g] = f.cs[c];
  f.omega[g] = f.omega[c];
  if (f.fCl) { f.fCl[g] = f.fCl[c]; f.fCq[g] = f.fCq[c]; f.fsh[g] = f.fsh[c]; }
}

__global__ void k_ghost_scalar(double* __restrict__ a, int64_t g0, int64_t g1, int64_t nInt,
                               const int32_t* __restrict__ control) {
  const int64_t g = g0 + blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (g < g1) a[g] = a[control[g - nInt]];
}

// A Vector field on ghost rows [g0, g1): R.a of the control for a reflecting
// plane, a copy for a periodic one (ReflectingBoundary.cc:106-125).
template <int D>
__global__ void k_ghost_vector(double* __restrict__ a, int64_t g0, int64_t g1, int64_t nInt,
                               const int32_t* __restrict__ control, DSub b) {
  const int64_t g = g0 + blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (g >= g1) return;
  const V<D> v = loadV<D>(a, control[g - nInt]);
  const V<D> o = b.reflect ? reflectV<D>(b.R, v) : v;
  for (int k = 0; k < D; ++k) a[g * D + k] = o.c[k];
}

// Ghost records after sum density: the same expressions as the internal update.
template <bool ISO>
__global__ void k_ghost_records(double* __restrict__ rec, const int32_t* __restrict__ iperm,
                                const double* __restrict__ rho, int64_t nInt, int64_t nTot) {
  using R = RL<ISO>;
  const int64_t g = nInt + blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (g >= nTot) return;
  double* r = rec + int64_t(iperm[g]) * R::STRIDE;
  const double acc = rho[g];
  const double som = r[R::SOM], p = r[R::P], m = r[R::M];
  r[R::RHO] = acc;
  r[R::PRHO] = som * p / (acc * acc);
  r[R::MOR] = m / acc;
}

// ---- boundary enforcement ---------------------------------------------------
// position < plane (GeomPlane::compare == 1): n.(x - p) < 0 and not within
// fuzzyEqual of 0 (src/Geometry/GeomPlane.cc:306-341).
template <int D> __device__ __forceinline__ bool behind(const double* P, const double* N, const V<D>& x) {
  const double d = signedDist<D>(P, N, x);
  return !fuzzyEqual(d, 0.0) && !(d > 0.0);
}

// PlanarBoundary::setViolationNodes / updateViolationNodes over the internal
// nodes, with ReflectingBoundary::enforceBoundary of v and H
// ((R*H)*R).Symmetric(), src/Boundary/ReflectingBoundary.cc:240-282).
template <int D>
__global__ void k_enforce(double* __restrict__ pos, double* __restrict__ vel, double* __restrict__ H, int64_t nInt,
                          DSub b) {
  constexpr int NS = NSym<D>::value;
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= nInt) return;
  const V<D> x = loadV<D>(pos, i);
  if (!behind<D>(b.enterP, b.enterN, x)) return;
  // mapPosition(x, enterPlane, exitPlane)
  const V<D> p0 = closestPoint<D>(b.exitP, b.exitN, x);
  const double sd = signedDist<D>(b.enterP, b.enterN, x);
  for (int k = 0; k < D; ++k) pos[i * D + k] = p0.c[k] - b.exitN[k] * sd;
  if (b.reflect) {
    const V<D> v = reflectV<D>(b.R, loadV<D>(vel, i));
    for (int k = 0; k < D; ++k) vel[i * D + k] = v.c[k];
    double full[D * D], t1[D * D], t2[D * D];
    int c = 0;
    for (int a = 0; a < D; ++a)
      for (int q = a; q < D; ++q, ++c) { full[a * D + q] = H[i * NS + c]; full[q * D + a] = H[i * NS + c]; }
    for (int a = 0; a < D; ++a)
      for (int q = 0; q < D; ++q) {
        double sum = b.R[a * D] * full[q];
        for (int k = 1; k < D; ++k) sum = sum + b.R[a * D + k] * full[k * D + q];
        t1[a * D + q] = sum;
      }
    for (int a = 0; a < D; ++a)
      for (int q = 0; q < D; ++q) {
        double sum = t1[a * D] * b.R[q];
        for (int k = 1; k < D; ++k) sum = sum + t1[a * D + k] * b.R[k * D + q];
        t2[a * D + q] = sum;
      }
    c = 0;
    for (int a = 0; a < D; ++a)
      for (int q = a; q < D; ++q, ++c) H[i * NS + c] = (a == q) ? t2[a * D + a] : 0.5 * (t2[a * D + q] + t2[q * D + a]);
  }
}

// PeriodicBoundary::updateViolationNodes: nodes still outside either plane
// are wrapped by the fmod rule (src/Boundary/PeriodicBoundary.cc:176-207).
template <int D>
__global__ void k_periodic_wrap(double* __restrict__ pos, int64_t nInt, DSub b) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= nInt) return;
  const V<D> x = loadV<D>(pos, i);
  // plane1 = enter, plane2 = exit of the first half of the pair
  if (!(behind<D>(b.enterP, b.enterN, x) || behind<D>(b.exitP, b.exitN, x))) return;
  const V<D> p2 = arr<D>(b.exitP);
  const double L = dabs(signedDist<D>(b.enterP, b.enterN, p2));
  const double s = vdot(vsub(arr<D>(b.enterP), x), arr<D>(b.enterN));
  const double f = dmax(-1.0, dmin(1.0, fmod(signedDist<D>(b.enterP, b.enterN, x) / L, 1.0)));
  for (int k = 0; k < D; ++k) pos[i * D + k] = (x.c[k] + b.enterN[k] * s) + b.enterN[k] * (f * L);
}

// ---- culling ---------------------------------------------------------------
// keep[c]: internal nodes, ghosts some row holds (marked by the search in slot
// space), then every ghost that controls another ghost.
__global__ void k_keep_gather(int32_t* __restrict__ keep, const int32_t* __restrict__ iperm,
                              const int32_t* __restrict__ slotKeep, int64_t nInt, int64_t nTot) {
  const int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (c < nInt) keep[c] = 1;
  else if (c < nTot) keep[c] = slotKeep[iperm[c]];
}

__global__ void k_keep_controls(int32_t* __restrict__ keep, int64_t nInt, int64_t nTot,
                                const int32_t* __restrict__ control) {
  const int64_t g = nInt + blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (g >= nTot) return;
  const int32_t c = control[g - nInt];
  if (c >= nInt) keep[c] = 1;
}

// Compact the ghost rows [i0, n) of one caller-order field of `comps`
// doubles per row (internal rows keep their place).
__global__ void k_compact_rows(const double* __restrict__ src, double* __restrict__ dst, int comps,
                               const int32_t* __restrict__ keep, const int32_t* __restrict__ scan, int64_t i0,
                               int64_t n) {
  const int64_t i = i0 + blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n || !keep[i]) return;
  const int64_t o = scan[i];
  for (int c = 0; c < comps; ++c) dst[o * comps + c] = src[i * comps + c];
}

__global__ void k_compact_ghosts(const int32_t* __restrict__ control, const int32_t* __restrict__ subOf,
                                 int32_t* __restrict__ newControl, int32_t* __restrict__ newSubOf,
                                 const int32_t* __restrict__ keep, const int32_t* __restrict__ scan,
                                 int64_t nInt, int64_t nTot) {
  const int64_t g = nInt + blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (g >= nTot || !keep[g]) return;
  const int64_t o = scan[g] - nInt;
  newControl[o] = scan[control[g - nInt]];
  newSubOf[o] = subOf[g - nInt];
}

// Slots keep their place; their caller index is renumbered, culled ones die.
// The records (when current) take the new caller index; nothing else in a
// record depends on the numbering, so they stay valid.
__global__ void k_perm_remap(int32_t* __restrict__ perm, const int32_t* __restrict__ keep,
                             const int32_t* __restrict__ scan, int64_t nSlots, double* __restrict__ rec, int stride,
                             int origOff) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k >= nSlots) return;
  const int32_t c = perm[k];
  if (c < 0) return;
  const int32_t nc = keep[c] ? scan[c] : -1;
  perm[k] = nc;
  const double o = nc >= 0 ? double(nc) : 2147483648.0;
  if (rec) rec[k * stride + origOff] = o;
}

__global__ void k_iperm_remap(const int32_t* __restrict__ iperm, const int32_t* __restrict__ keep,
                              const int32_t* __restrict__ scan, int64_t n, int32_t* __restrict__ out) {
  const int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (c < n && keep[c]) out[scan[c]] = iperm[c];
}

std::vector<DBuf*> stateBufs(Sim& s) {
  return {&s.pos, &s.vel, &s.H, &s.mass, &s.rho, &s.eps, &s.P, &s.cs, &s.omega, &s.fCl, &s.fCq, &s.fshear, &s.seg};
}
// fields 9-11 exist with Q multipliers, field 12 in an ensemble
bool stateUsed(const Sim& s, int which) {
  if (which >= 9 && which < 12) return s.hasMultipliers;
  if (which == 12) return s.ensemble || s.nLists > 1;
  return true;
}
int stateComps(const Sim& s, int which) {
  const int D = s.dim, NS = nsymOf(D);
  return which == 0 || which == 1 ? D : (which == 2 ? NS : 1);
}

// Make room for `rows` rows in every caller-order field, keeping nTot rows.
void ensureRows(Sim& s, int64_t rows) {
  auto bufs = stateBufs(s);
  for (int w = 0; w < int(bufs.size()); ++w) {
    if (!stateUsed(s, w)) continue;
    const size_t per = sizeof(double) * size_t(stateComps(s, w));
    if (bufs[size_t(w)]->bytes >= per * size_t(rows)) continue;
    bufs[size_t(w)]->grow(per * size_t(rows + rows / 4), per * size_t(s.nTot), s.stream);
  }
}

FieldPtrs fieldPtrs(Sim& s) {
  FieldPtrs f;
  f.pos = s.pos.as<double>(); f.H = s.H.as<double>(); f.vel = s.vel.as<double>(); f.mass = s.mass.as<double>();
  f.rho = s.rho.as<double>(); f.eps = s.eps.as<double>(); f.P = s.P.as<double>(); f.cs = s.cs.as<double>();
  f.omega = s.omega.as<double>();
  f.fCl = s.hasMultipliers ? s.fCl.as<double>() : nullptr;
  f.fCq = s.hasMultipliers ? s.fCq.as<double>() : nullptr;
  f.fsh = s.hasMultipliers ? s.fshear.as<double>() : nullptr;
  return f;
}

template <int D> void applyFieldsD(Sim& s, bool geometry) {
  const FieldPtrs f = fieldPtrs(s);
  for (size_t b = 0; b < s.subs.size(); ++b) {
    const int64_t g0 = s.nInt + s.subStart[b], g1 = s.nInt + s.subStart[b + 1];
    if (g1 <= g0) continue;
    hipLaunchKernelGGL(k_ghost_fields<D>, dim3(nblocks(g1 - g0)), dim3(kBlock), 0, s.stream, f, g0, g1, s.nInt,
                       s.ghostControl.as<int32_t>(), toDevice(s.subs[b]), geometry ? 1 : 0);
  }
  SG_CHECK_HIP(hipGetLastError());
}

// Every plane's ghosts with the running row count kept on the device: the
// grids cover a capacity guessed from the previous generation, and one
// readback at the end returns the count and each sub-boundary's first ghost.
// Returns false when a plane's ghosts would pass the capacity (the caller then
// generates again plane by plane; generation reads only internal rows).
template <int D> bool generateFastD(Sim& s, double ext, unsigned long long* hbits, int nh) {
  hipStream_t st = s.stream;
  const int64_t nInt = s.nInt;
  const int64_t est = s.ghostsGenerated + s.ghostsGenerated / 4 + 1024;
  const int64_t cap = nInt + est;
  SG_REQUIRE(cap < (int64_t(1) << 31) - 1, SPHGPU_ERR_CAPACITY, "too many ghost nodes");
  ensureRows(s, cap);
  s.hmaxOf.grow(sizeof(double) * size_t(cap), sizeof(double) * size_t(nInt), st);
  s.ghostControl.grow(sizeof(int32_t) * size_t(est), 0, st);
  s.ghostSub.grow(sizeof(int32_t) * size_t(est), 0, st);
  s.keep.ensure(sizeof(int32_t) * size_t(cap + 1));
  // per-workgroup flag counts (at nb + 1) and their exclusive scan (at 0)
  const int64_t nb = nblocks(cap);
  s.keepScan.ensure(sizeof(int32_t) * size_t(std::max<int64_t>(cap + 1, 2 * (nb + 1))));
  int32_t* blkScan = s.keepScan.as<int32_t>();
  int32_t* blkCount = blkScan + (nb + 1);
  s.genDev.ensure(sizeof(GenDev));
  GenDev* g = s.genDev.as<GenDev>();
  const int64_t* nDev = &g->n;
  hipLaunchKernelGGL(k_gen_init, dim3(1), dim3(kBlock), 0, st, g, nInt, hbits, nh);
  const double* segp = groupCount(s) > 0 ? s.seg.as<double>() : nullptr;
  for (size_t b = 0; b < s.subs.size(); ++b) {
    const DSub d = toDevice(s.subs[b]);
    if (segp)
      hipLaunchKernelGGL(k_ghost_hmax<D>, dim3(nblocks(cap)), dim3(kBlock), 0, st, s.pos.as<double>(),
                         s.hmaxOf.as<double>(), segp, cap, d, ext, hbits, nDev);
    else
      hipLaunchKernelGGL(k_ghost_hmax_one<D>, dim3(std::min(nblocks(cap), 1024)), dim3(kBlock), 0, st,
                         s.pos.as<double>(), s.hmaxOf.as<double>(), cap, d, ext, hbits, nDev);
    hipLaunchKernelGGL(k_ghost_flag<D>, dim3(nb), dim3(kBlock), 0, st, s.pos.as<double>(), segp, cap, d, ext, hbits,
                       s.keep.as<int32_t>(), nDev, blkCount);
    exclusiveScanI32(blkCount, blkScan, nb, s);
    hipLaunchKernelGGL((k_ghost_make<D, true>), dim3(nb), dim3(kBlock), 0, st, s.pos.as<double>(), s.H.as<double>(),
                       groupCount(s) > 0 ? s.seg.as<double>() : nullptr, cap, nInt, s.keep.as<int32_t>(), blkScan, d, int(b),
                       s.ghostControl.as<int32_t>(), s.ghostSub.as<int32_t>(), s.hmaxOf.as<double>(), nDev, cap,
                       &g->overflow);
    hipLaunchKernelGGL(k_gen_advance, dim3(1), dim3(kBlock), 0, st, g, int(b), blkScan + nb, nInt, cap, hbits, nh);
    SG_CHECK_HIP(hipGetLastError());
  }
  GenDev* h = static_cast<GenDev*>(hostStaging(s, 0, sizeof(GenDev)));
  SG_CHECK_HIP(hipMemcpyAsync(h, g, sizeof(GenDev), hipMemcpyDeviceToHost, st));
  SG_CHECK_HIP(hipStreamSynchronize(st));
  if (h->overflow) return false;
  s.nTot = h->n;
  for (size_t b = 0; b < s.subs.size(); ++b) s.subStart[b] = h->start[b];
  return true;
}

template <int D> void generateD(Sim& s, double ext) {
  hipStream_t st = s.stream;
  constexpr int NS = NSym<D>::value;
  (void)NS;
  s.subStart.assign(s.subs.size() + 1, 0);
  s.flags.ensure(sizeof(int32_t) * 8 + sizeof(unsigned long long) * 2);
  unsigned long long* hbits = reinterpret_cast<unsigned long long*>(s.flags.as<int32_t>() + 4);
  const int nh = std::max(1, groupCount(s));
  if (groupCount(s) > 0) {
    s.segHmaxBits.ensure(sizeof(unsigned long long) * size_t(nh));
    hbits = s.segHmaxBits.as<unsigned long long>();
  }
  s.hmaxOf.ensure(sizeof(double) * size_t(std::max<int64_t>(s.nInt, 1)));
  if (s.nInt > 0)
    hipLaunchKernelGGL(k_hinv_min<D>, dim3(nblocks(s.nInt)), dim3(kBlock), 0, st, s.H.as<double>(), s.nInt,
                       s.hmaxOf.as<double>());
  // after a first generation, one readback for all planes
  bool done = false;
  if (s.ghostsGenerated > 0 && !s.subs.empty() && s.subs.size() < 16) {
    done = generateFastD<D>(s, ext, hbits, nh);
    if (!done) s.nTot = s.nInt;
  }
  for (size_t b = 0; !done && b < s.subs.size(); ++b) {
    const int64_t n = s.nTot;
    s.subStart[b] = n - s.nInt;
    const DSub d = toDevice(s.subs[b]);
    s.keep.ensure(sizeof(int32_t) * size_t(n + 1));
    s.keepScan.ensure(sizeof(int32_t) * size_t(n + 1));
    const double* segp = groupCount(s) > 0 ? s.seg.as<double>() : nullptr;
    SG_CHECK_HIP(hipMemsetAsync(hbits, 0, sizeof(unsigned long long) * size_t(nh), st));
    if (segp)
      hipLaunchKernelGGL(k_ghost_hmax<D>, dim3(nblocks(n)), dim3(kBlock), 0, st, s.pos.as<double>(),
                         s.hmaxOf.as<double>(), segp, n, d, ext, hbits, nullptr);
    else
      hipLaunchKernelGGL(k_ghost_hmax_one<D>, dim3(std::min(nblocks(n), 1024)), dim3(kBlock), 0, st,
                         s.pos.as<double>(), s.hmaxOf.as<double>(), n, d, ext, hbits, nullptr);
    hipLaunchKernelGGL(k_ghost_flag<D>, dim3(nblocks(n)), dim3(kBlock), 0, st, s.pos.as<double>(), segp, n, d, ext,
                       hbits, s.keep.as<int32_t>(), nullptr, nullptr);
    exclusiveScanI32(s.keep.as<int32_t>(), s.keepScan.as<int32_t>(), n, s);
    int32_t* hc = static_cast<int32_t*>(hostStaging(s, 0, sizeof(int32_t)));
    SG_CHECK_HIP(hipMemcpyAsync(hc, s.keepScan.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SG_CHECK_HIP(hipStreamSynchronize(st));
    const int32_t count = *hc;
    SG_REQUIRE(n + int64_t(count) < (int64_t(1) << 31) - 1, SPHGPU_ERR_CAPACITY, "too many ghost nodes");
    if (count == 0) continue;
    ensureRows(s, n + count);
    s.hmaxOf.grow(sizeof(double) * size_t(n + count + (n + count) / 4), sizeof(double) * size_t(n), st);
    const int64_t ng = n + count - s.nInt;
    s.ghostControl.grow(sizeof(int32_t) * size_t(ng + ng / 2), sizeof(int32_t) * size_t(n - s.nInt), st);
    s.ghostSub.grow(sizeof(int32_t) * size_t(ng + ng / 2), sizeof(int32_t) * size_t(n - s.nInt), st);
    hipLaunchKernelGGL((k_ghost_make<D, false>), dim3(nblocks(n)), dim3(kBlock), 0, st, s.pos.as<double>(), s.H.as<double>(),
                       groupCount(s) > 0 ? s.seg.as<double>() : nullptr, n, s.nInt, s.keep.as<int32_t>(), s.keepScan.as<int32_t>(), d, int(b),
                       s.ghostControl.as<int32_t>(), s.ghostSub.as<int32_t>(), s.hmaxOf.as<double>(), nullptr,
                       int64_t(0), nullptr);
    SG_CHECK_HIP(hipGetLastError());
    s.nTot = n + count;
  }
  s.subStart[s.subs.size()] = s.nTot - s.nInt;
  s.ghostsGenerated = s.nTot - s.nInt;
  applyFieldsD<D>(s, false);
}

}  // namespace

// Integrator::cullGhostNodes, first half: keep flags (ghosts some internal row
// uses, and the controls they chain through), their scan, and the new total
// and first ghost of every sub-boundary copied to pinned `hostOut` (nsub + 1
// values, read after the caller's next wait).  The neighbour build issues it
// before its row readback so that one wait serves both.
void cullLaunch(Sim& s, int32_t* hostOut) {
  hipStream_t st = s.stream;
  const int64_t n = s.nTot, nInt = s.nInt;
  const size_t nsub = s.subStart.size();
  SG_REQUIRE(n > nInt && nsub < 16, SPHGPU_ERR_STATE, "cullLaunch needs ghosts of at most 15 sub-boundaries");
  s.keep.ensure(sizeof(int32_t) * size_t(n + 1));
  s.keepScan.ensure(sizeof(int32_t) * size_t(n + 1));
  int32_t* keep = s.keep.as<int32_t>();
  hipLaunchKernelGGL(k_keep_gather, dim3(nblocks(n)), dim3(kBlock), 0, st, keep, s.iperm.as<int32_t>(),
                     s.slotKeep.as<int32_t>(), nInt, n);
  hipLaunchKernelGGL(k_keep_controls, dim3(nblocks(n - nInt)), dim3(kBlock), 0, st, keep, nInt, n,
                     s.ghostControl.as<int32_t>());
  exclusiveScanI32(keep, s.keepScan.as<int32_t>(), n, s);
  int64_t at[16];
  for (size_t b = 0; b < nsub; ++b) at[b] = nInt + s.subStart[b];
  at[nsub] = n;
  pickInts(s, s.keepScan.as<int32_t>(), at, int(nsub + 1), hostOut);
}

namespace {

void cull(Sim& s) {
  hipStream_t st = s.stream;
  const int64_t n = s.nTot, nInt = s.nInt;
  if (n == nInt) return;
  const size_t nsub = s.subStart.size();
  std::vector<int32_t> starts(nsub + 1);
  if (s.cullStartsReady) {
    // launched and read with the neighbour build's row readback
    SG_REQUIRE(s.cullStarts.size() == nsub + 1, SPHGPU_ERR_STATE, "stale cull starts");
    starts = s.cullStarts;
    s.cullStartsReady = false;
  } else if (nsub < 16) {
    int32_t* h = static_cast<int32_t*>(hostStaging(s, 0, sizeof(int32_t) * 16));
    cullLaunch(s, h);
    SG_CHECK_HIP(hipStreamSynchronize(st));
    std::copy(h, h + nsub + 1, starts.begin());
  } else {
    s.keep.ensure(sizeof(int32_t) * size_t(n + 1));
    s.keepScan.ensure(sizeof(int32_t) * size_t(n + 1));
    int32_t* keep = s.keep.as<int32_t>();
    hipLaunchKernelGGL(k_keep_gather, dim3(nblocks(n)), dim3(kBlock), 0, st, keep, s.iperm.as<int32_t>(),
                       s.slotKeep.as<int32_t>(), nInt, n);
    hipLaunchKernelGGL(k_keep_controls, dim3(nblocks(n - nInt)), dim3(kBlock), 0, st, keep, nInt, n,
                       s.ghostControl.as<int32_t>());
    exclusiveScanI32(keep, s.keepScan.as<int32_t>(), n, s);
    int32_t* h = static_cast<int32_t*>(hostStaging(s, 0, sizeof(int32_t) * (nsub + 1)));
    for (size_t b = 0; b < nsub; ++b)
      SG_CHECK_HIP(hipMemcpyAsync(h + b, s.keepScan.as<int32_t>() + nInt + s.subStart[b], sizeof(int32_t),
                                  hipMemcpyDeviceToHost, st));
    SG_CHECK_HIP(hipMemcpyAsync(h + nsub, s.keepScan.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SG_CHECK_HIP(hipStreamSynchronize(st));
    std::copy(h, h + nsub + 1, starts.begin());
  }
  int32_t* keep = s.keep.as<int32_t>();
  const int64_t nNew = starts.back();
  if (nNew == n) return;
  const int32_t* scan = s.keepScan.as<int32_t>();

  // caller-order fields: kept ghost rows compacted through the spare and
  // copied back (internal rows stay where they are)
  auto bufs = stateBufs(s);
  for (int w = 0; w < int(bufs.size()); ++w) {
    if (!stateUsed(s, w)) continue;
    const int comps = stateComps(s, w);
    DBuf& spare = s.spare[w];
    spare.ensure(bufs[size_t(w)]->bytes);
    hipLaunchKernelGGL(k_compact_rows, dim3(nblocks(n - nInt)), dim3(kBlock), 0, st, bufs[size_t(w)]->as<double>(),
                       spare.as<double>(), comps, keep, scan, nInt, n);
    if (nNew > nInt)
      SG_CHECK_HIP(hipMemcpyAsync(bufs[size_t(w)]->as<double>() + nInt * comps, spare.as<double>() + nInt * comps,
                                  sizeof(double) * size_t(nNew - nInt) * comps, hipMemcpyDeviceToDevice, st));
  }
  // ghost controls (control and sub-boundary arrays share the spare pair)
  s.ghostControlSpare.ensure(s.ghostControl.bytes);
  s.ghostSubSpare.ensure(s.ghostSub.bytes);
  hipLaunchKernelGGL(k_compact_ghosts, dim3(nblocks(n - nInt)), dim3(kBlock), 0, st, s.ghostControl.as<int32_t>(),
                     s.ghostSub.as<int32_t>(), s.ghostControlSpare.as<int32_t>(), s.ghostSubSpare.as<int32_t>(), keep,
                     scan, nInt, n);
  s.ghostControl.swap(s.ghostControlSpare);
  s.ghostSub.swap(s.ghostSubSpare);
  // slot <-> caller maps
  const bool recs = s.recValid;
  // The density records keep the pre-cull caller index of their ghosts: the
  // density sum compares a neighbour's index only with an internal node's
  // (to place the self term), and culling keeps every internal index and
  // every ghost's index above them, so no comparison changes.
  hipLaunchKernelGGL(k_perm_remap, dim3(nblocks(s.nSlots)), dim3(kBlock), 0, st, s.perm.as<int32_t>(), keep, scan,
                     s.nSlots, recs ? s.rec.as<double>() : nullptr, s.isotropic ? RL<true>::STRIDE : RL<false>::STRIDE,
                     s.isotropic ? RL<true>::ORIG : RL<false>::ORIG);
  s.ipermSpare.ensure(s.iperm.bytes);
  hipLaunchKernelGGL(k_iperm_remap, dim3(nblocks(n)), dim3(kBlock), 0, st, s.iperm.as<int32_t>(), keep, scan, n,
                     s.ipermSpare.as<int32_t>());
  s.iperm.swap(s.ipermSpare);
  SG_CHECK_HIP(hipGetLastError());
  for (size_t b = 0; b < s.subStart.size(); ++b) s.subStart[b] = int64_t(starts[b]) - nInt;
  s.nTot = nNew;
  s.recValid = recs;
  // with several NodeLists the density sum also looks up a neighbour's
  // NodeList by that index: the density records are packed again
  if (s.nLists > 1) s.recSValid = false;
  s.derivsValid = false;
}

}  // namespace

void setBoundaries(Sim& s, int n, const sphgpu_boundary* b) {
  SG_REQUIRE(n >= 0 && (n == 0 || b), SPHGPU_ERR_ARG, "bad boundary list");
  const int D = s.dim;
  std::vector<SubBoundary> subs;
  auto unit = [&](const double* nin, double* nout) {
    // GeomPlane: normal.unitVector() (1-D divides, else multiplies by 1/|n|)
    double m2 = nin[0] * nin[0];
    for (int k = 1; k < D; ++k) m2 = m2 + nin[k] * nin[k];
    const double m = D == 1 ? std::fabs(nin[0]) : std::sqrt(m2);
    SG_REQUIRE(m > 1.0e-50, SPHGPU_ERR_ARG, "boundary plane normal is zero");
    for (int k = 0; k < 3; ++k) nout[k] = 0.0;
    if (D == 1) nout[0] = nin[0] / m;
    else { const double inv = 1.0 / m; for (int k = 0; k < D; ++k) nout[k] = nin[k] * inv; }
  };
  for (int q = 0; q < n; ++q) {
    SG_REQUIRE(b[q].kind == SPHGPU_BOUNDARY_REFLECTING || b[q].kind == SPHGPU_BOUNDARY_PERIODIC, SPHGPU_ERR_ARG,
               "unknown boundary kind");
    double n1[3], n2[3], p1[3] = {0, 0, 0}, p2[3] = {0, 0, 0};
    unit(b[q].plane1.normal, n1);
    for (int k = 0; k < D; ++k) p1[k] = b[q].plane1.point[k];
    SubBoundary a{};
    if (b[q].kind == SPHGPU_BOUNDARY_REFLECTING) {
      for (int k = 0; k < 3; ++k) { a.enterP[k] = a.exitP[k] = p1[k]; a.enterN[k] = a.exitN[k] = n1[k]; }
      a.reflect = 1;
      a.periodicSecond = 0;
      // Tensor::one - 2.0*normal.selfdyad()
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) a.R[i * D + j] = (i == j ? 1.0 : 0.0) - 2.0 * (n1[std::min(i, j)] * n1[std::max(i, j)]);
      subs.push_back(a);
    } else {
      unit(b[q].plane2.normal, n2);
      for (int k = 0; k < D; ++k) p2[k] = b[q].plane2.point[k];
      SubBoundary c{}, d{};
      for (int k = 0; k < 3; ++k) {
        c.enterP[k] = p1[k]; c.enterN[k] = n1[k]; c.exitP[k] = p2[k]; c.exitN[k] = n2[k];
        d.enterP[k] = p2[k]; d.enterN[k] = n2[k]; d.exitP[k] = p1[k]; d.exitN[k] = n1[k];
      }
      c.reflect = d.reflect = 0;
      c.periodicSecond = 0;
      d.periodicSecond = 1;
      subs.push_back(c);
      subs.push_back(d);
    }
  }
  s.subs = subs;
  s.subStart.clear();
}

void setGhostNodes(Sim& s, double ext, bool doCull) {
  SG_REQUIRE(ext > 0.0, SPHGPU_ERR_ARG, "kernel extent must be positive");
  SG_REQUIRE(s.nTot == s.nInt || s.deviceGhosts, SPHGPU_ERR_STATE,
             "sphgpu_set_ghost_nodes needs n_total = n_internal (ghosts are made on the device)");
  hipStream_t st = s.stream;
  SG_CHECK_HIP(hipEventRecord(s.ev[8], st));
  s.nTot = s.nInt;                       // drop the previous ghosts
  s.neighborsValid = false;
  s.recValid = false;
  s.derivsValid = false;
  if (s.dim == 1) generateD<1>(s, ext);
  else if (s.dim == 2) generateD<2>(s, ext);
  else generateD<3>(s, ext);
  s.deviceGhosts = true;
  SG_CHECK_HIP(hipEventRecord(s.ev[9], st));
  s.markGhostUse = doCull;
  try {
    buildNeighbors(s, ext);
  } catch (...) {
    s.markGhostUse = false;
    throw;
  }
  s.markGhostUse = false;
  SG_CHECK_HIP(hipEventRecord(s.ev[10], st));
  if (doCull) cull(s);
  SG_CHECK_HIP(hipEventRecord(s.ev[11], st));
  s.timingPending |= 8u;
}

void applyGhostBoundaries(Sim& s) {
  if (!s.deviceGhosts || s.nTot == s.nInt) return;
  if (s.dim == 1) applyFieldsD<1>(s, true);
  else if (s.dim == 2) applyFieldsD<2>(s, true);
  else applyFieldsD<3>(s, true);
  s.recValid = false;
  s.derivsValid = false;
}

void enforceBoundaries(Sim& s) {
  if (s.subs.empty() || s.nInt == 0) return;
  hipStream_t st = s.stream;
  for (size_t b = 0; b < s.subs.size(); ++b) {
    const DSub d = toDevice(s.subs[b]);
    if (s.dim == 1)
      hipLaunchKernelGGL(k_enforce<1>, dim3(nblocks(s.nInt)), dim3(kBlock), 0, st, s.pos.as<double>(), s.vel.as<double>(),
                         s.H.as<double>(), s.nInt, d);
    else if (s.dim == 2)
      hipLaunchKernelGGL(k_enforce<2>, dim3(nblocks(s.nInt)), dim3(kBlock), 0, st, s.pos.as<double>(), s.vel.as<double>(),
                         s.H.as<double>(), s.nInt, d);
    else
      hipLaunchKernelGGL(k_enforce<3>, dim3(nblocks(s.nInt)), dim3(kBlock), 0, st, s.pos.as<double>(), s.vel.as<double>(),
                         s.H.as<double>(), s.nInt, d);
    // the second half of a periodic pair closes it: wrap what is still outside
    if (s.subs[b].periodicSecond) {
      const DSub first = toDevice(s.subs[b - 1]);
      if (s.dim == 1)
        hipLaunchKernelGGL(k_periodic_wrap<1>, dim3(nblocks(s.nInt)), dim3(kBlock), 0, st, s.pos.as<double>(), s.nInt, first);
      else if (s.dim == 2)
        hipLaunchKernelGGL(k_periodic_wrap<2>, dim3(nblocks(s.nInt)), dim3(kBlock), 0, st, s.pos.as<double>(), s.nInt, first);
      else
        hipLaunchKernelGGL(k_periodic_wrap<3>, dim3(nblocks(s.nInt)), dim3(kBlock), 0, st, s.pos.as<double>(), s.nInt, first);
    }
  }
  SG_CHECK_HIP(hipGetLastError());
  s.recValid = false;
  s.derivsValid = false;
}

void applyGhostScalar(Sim& s, DBuf& field) {
  if (!s.deviceGhosts || s.nTot == s.nInt) return;
  for (size_t b = 0; b < s.subs.size(); ++b) {
    const int64_t g0 = s.nInt + s.subStart[b], g1 = s.nInt + s.subStart[b + 1];
    if (g1 <= g0) continue;
    hipLaunchKernelGGL(k_ghost_scalar, dim3(nblocks(g1 - g0)), dim3(kBlock), 0, s.stream, field.as<double>(), g0, g1,
                       s.nInt, s.ghostControl.as<int32_t>());
  }
  SG_CHECK_HIP(hipGetLastError());
}

void ensureStateRows(Sim& s, int64_t rows) { ensureRows(s, rows); }
bool stateFieldUsed(const Sim& s, int which) { return stateUsed(s, which); }

namespace {
__global__ void k_fill_value(double* __restrict__ a, int64_t first, int64_t rows, double v) {
  const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (r < rows) a[first + r] = v;
}
}  // namespace

// Several NodeLists in one node set: NodeList L owns `counts[L]` consecutive
// internal rows, in NodeList order; their ghosts are made by the device
// boundaries, each NodeList with its own (PlanarBoundary::setGhostNodes per
// NodeList).  gamma (may be null) gives each NodeList its gamma-law EOS.
void setNodeLists(Sim& s, int n, const int64_t* counts, const double* gamma, int sumAll) {
  SG_REQUIRE(!s.ensemble, SPHGPU_ERR_STATE, "an ensemble holds independent simulations, not NodeLists");
  SG_REQUIRE(n >= 1 && counts, SPHGPU_ERR_ARG, "need at least one NodeList");
  SG_REQUIRE(s.nTot == s.nInt, SPHGPU_ERR_STATE,
             "set the NodeLists on internal rows only (n_total = n_internal); the device makes their ghosts");
  std::vector<int64_t> start(size_t(n) + 1, 0);
  for (int L = 0; L < n; ++L) {
    SG_REQUIRE(counts[L] >= 0, SPHGPU_ERR_ARG, "negative NodeList size");
    start[size_t(L) + 1] = start[size_t(L)] + counts[L];
  }
  SG_REQUIRE(start.back() == s.nInt, SPHGPU_ERR_ARG, "the NodeList sizes must add up to n_internal");
  if (gamma)
    for (int L = 0; L < n; ++L) SG_REQUIRE(gamma[L] > 1.0, SPHGPU_ERR_ARG, "a NodeList gamma must exceed 1");
  s.nLists = n > 1 ? n : 0;
  s.listStart = start;
  s.sumAllLists = sumAll != 0;
  s.listGammaHost.assign(gamma ? gamma : nullptr, gamma ? gamma + n : nullptr);
  if (gamma) {
    s.listGamma.ensure(sizeof(double) * size_t(n));
    SG_CHECK_HIP(hipMemcpyAsync(s.listGamma.p, gamma, sizeof(double) * size_t(n), hipMemcpyHostToDevice, s.stream));
  }
  if (s.nLists > 1) {
    ensureRows(s, std::max<int64_t>(s.nTot, 1));
    for (int L = 0; L < n; ++L)
      if (counts[L] > 0)
        hipLaunchKernelGGL(k_fill_value, dim3(nblocks(counts[L])), dim3(kBlock), 0, s.stream, s.seg.as<double>(),
                           start[size_t(L)], counts[L], double(L));
    SG_CHECK_HIP(hipGetLastError());
  }
  SG_CHECK_HIP(hipStreamSynchronize(s.stream));   // `gamma` may be released on return
  s.neighborsValid = false;
  s.recValid = false;
  s.derivsValid = false;
}
DBuf& stateField(Sim& s, int which) { return *stateBufs(s)[size_t(which)]; }
int stateFieldComps(const Sim& s, int which) { return stateComps(s, which); }

void applyGhostVector(Sim& s, DBuf& field) {
  if (!s.deviceGhosts || s.nTot == s.nInt) return;
  for (size_t b = 0; b < s.subs.size(); ++b) {
    const int64_t g0 = s.nInt + s.subStart[b], g1 = s.nInt + s.subStart[b + 1];
    if (g1 <= g0) continue;
    const DSub d = toDevice(s.subs[b]);
    const int32_t* ctl = s.ghostControl.as<int32_t>();
    if (s.dim == 1)
      hipLaunchKernelGGL(k_ghost_vector<1>, dim3(nblocks(g1 - g0)), dim3(kBlock), 0, s.stream, field.as<double>(), g0,
                         g1, s.nInt, ctl, d);
    else if (s.dim == 2)
      hipLaunchKernelGGL(k_ghost_vector<2>, dim3(nblocks(g1 - g0)), dim3(kBlock), 0, s.stream, field.as<double>(), g0,
                         g1, s.nInt, ctl, d);
    else
      hipLaunchKernelGGL(k_ghost_vector<3>, dim3(nblocks(g1 - g0)), dim3(kBlock), 0, s.stream, field.as<double>(), g0,
                         g1, s.nInt, ctl, d);
  }
  SG_CHECK_HIP(hipGetLastError());
}

void applyGhostDensity(Sim& s) {
  if (!s.deviceGhosts || s.nTot == s.nInt) return;
  hipStream_t st = s.stream;
  for (size_t b = 0; b < s.subs.size(); ++b) {
    const int64_t g0 = s.nInt + s.subStart[b], g1 = s.nInt + s.subStart[b + 1];
    if (g1 <= g0) continue;
    hipLaunchKernelGGL(k_ghost_scalar, dim3(nblocks(g1 - g0)), dim3(kBlock), 0, st, s.rho.as<double>(), g0, g1,
                       s.nInt, s.ghostControl.as<int32_t>());
  }
  const int64_t ng = s.nTot - s.nInt;
  if (s.isotropic)
    hipLaunchKernelGGL(k_ghost_records<true>, dim3(nblocks(ng)), dim3(kBlock), 0, st, s.rec.as<double>(),
                       s.iperm.as<int32_t>(), s.rho.as<double>(), s.nInt, s.nTot);
  else
    hipLaunchKernelGGL(k_ghost_records<false>, dim3(nblocks(ng)), dim3(kBlock), 0, st, s.rec.as<double>(),
                       s.iperm.as<int32_t>(), s.rho.as<double>(), s.nInt, s.nTot);
  SG_CHECK_HIP(hipGetLastError());
}

}  // namespace sphgpu
