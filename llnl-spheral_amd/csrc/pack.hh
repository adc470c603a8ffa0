// Record packing shared by the record pack (hydro.hip) and the integrator
// update (integrate.hip), so that both write bitwise the same records.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>
#include "dgeom.hh"
#include "records.hh"

namespace sphgpu {

// ---------------------------------------------------------------------------
// Records
// ---------------------------------------------------------------------------
// The neighbour records of the isotropic pair and density kernels, each made
// from an FP64 record r (isotropic layout).  The kernels' time is set by the
// number of 16-B loads per neighbour gather, so each record holds only what
// its kernel reads.
// The mixed pair kernel's record (FP32), float4s per slot by dimension: 1-D
// (x, +-1/h, m | 1/rho, v, c, P/rho^2/omega), 2-D (x, y | +-1/h, m, 1/rho, vx |
// vy, c, P/rho^2/omega, m/rho) with x, y as FP64 bits, 3-D (qx, qy, qz, +-1/h |
// m, 1/rho, vx, vy | vz, c, P/rho^2/omega, m/rho).  1/h is negative for ghost
// rows; det H = (1/h)^D and in 1-D m/rho = m * (1/rho) are formed by the reader.
// In 3-D the position is the low 32 bits of round(x / u) for a power of two u
// with the largest search radius below 2^30 u: the 32-bit difference of two
// neighbours' words is the exact difference of their rounded positions, so
// r_ij carries an error of at most u (1e-9 of the search radius), far below
// the FP32 rounding of r_ij itself.
__device__ __forceinline__ uint32_t quantPos(double x, double qInv) {
  return uint32_t(uint64_t(static_cast<long long>(rint(x * qInv))));
}
template <int D> struct MixedRec {
  static constexpr int F4 = D == 1 ? 2 : 3;
};
__device__ __forceinline__ float dlo(double x) { return __int_as_float(__double2loint(x)); }
__device__ __forceinline__ float dhi(double x) { return __int_as_float(__double2hiint(x)); }
template <int D>
__device__ __forceinline__ void writeMixedRec(const double* r, int64_t nInt, double qInv, float4* __restrict__ p) {
  using R = RL<true>;
  const bool ghost = r[R::ORIG] >= double(nInt);
  const float hinv = ghost ? -float(r[R::HINV]) : float(r[R::HINV]);
  const double rho = r[R::RHO];
  const float invRho = rho != 0.0 ? float(1.0 / rho) : 0.0f;
  const float m = float(r[R::M]), cs = float(r[R::CS]), prho = float(r[R::PRHO]), mor = float(r[R::MOR]);
  if constexpr (D == 1) {
    p[0] = make_float4(dlo(r[R::X]), dhi(r[R::X]), hinv, m);
    p[1] = make_float4(invRho, float(r[R::V]), cs, prho);
  } else if constexpr (D == 2) {
    p[0] = make_float4(dlo(r[R::X]), dhi(r[R::X]), dlo(r[R::X + 1]), dhi(r[R::X + 1]));
    p[1] = make_float4(hinv, m, invRho, float(r[R::V]));
    p[2] = make_float4(float(r[R::V + 1]), cs, prho, mor);
  } else {
    p[0] = make_float4(__uint_as_float(quantPos(r[R::X], qInv)), __uint_as_float(quantPos(r[R::X + 1], qInv)),
                       __uint_as_float(quantPos(r[R::X + 2], qInv)), hinv);
    p[1] = make_float4(m, invRho, float(r[R::V]), float(r[R::V + 1]));
    p[2] = make_float4(float(r[R::V + 2]), cs, prho, mor);
  }
}
// The FP64 isotropic pair kernel's record, 96 B per slot in two planes so that
// the gathered arrays (230 MB at 2.4 M slots) fit the 256 MB Infinity Cache:
// plane a (64 B): x y z, 1/h (negative for ghosts), m, 1/rho, c, P/rho^2/omega;
// plane b (32 B): v, 0.  det H and m/rho are recomputed by the loader.
template <int D>
__device__ __forceinline__ void writeFast64Rec(const double* r, int64_t nInt, double2* __restrict__ a,
                                               double2* __restrict__ b) {
  using R = RL<true>;
  const bool ghost = r[R::ORIG] >= double(nInt);
  const double rho = r[R::RHO];
  a[0] = make_double2(r[R::X], D > 1 ? r[R::X + 1] : 0.0);
  a[1] = make_double2(D > 2 ? r[R::X + 2] : 0.0, ghost ? -r[R::HINV] : r[R::HINV]);
  a[2] = make_double2(r[R::M], rho != 0.0 ? 1.0 / rho : 0.0);
  a[3] = make_double2(r[R::CS], r[R::PRHO]);
  b[0] = make_double2(r[R::V], D > 1 ? r[R::V + 1] : 0.0);
  b[1] = make_double2(D > 2 ? r[R::V + 2] : 0.0, 0.0);
}

// The density sum's neighbour record (DensRec).
template <int D>
__device__ __forceinline__ void writeDensRec(const double* r, double2* __restrict__ p) {
  using R = RL<true>;
  p[0] = make_double2(r[R::X], D > 1 ? r[R::X + 1] : 0.0);
  p[1] = make_double2(D > 2 ? r[R::X + 2] : 0.0, r[R::HINV]);
  p[2] = make_double2(r[R::M], r[R::ORIG]);
}

// The mixed density sum's neighbour record: x (FP64 bits), 1/h and m (FP32),
// one float4 in 1-D, two otherwise; det H = (1/h)^D is formed by the reader.
template <int D> struct DensMixedRec {
  static constexpr int F4 = D == 1 ? 1 : 2;
};
template <int D>
__device__ __forceinline__ void writeDensRecMixed(const double* r, float4* __restrict__ p) {
  using R = RL<true>;
  if constexpr (D == 1) {
    p[0] = make_float4(dlo(r[R::X]), dhi(r[R::X]), float(r[R::HINV]), float(r[R::M]));
  } else {
    const double x1 = r[R::X + 1], x2 = D > 2 ? r[R::X + 2] : 0.0;
    p[0] = make_float4(dlo(r[R::X]), dhi(r[R::X]), dlo(x1), dhi(x1));
    p[1] = make_float4(dlo(x2), dhi(x2), float(r[R::HINV]), float(r[R::M]));
  }
}

// The caller-order fields a record is made from.
struct PackIn {
  const int32_t* perm;     // slot -> caller row (-1: dead slot)
  const double *pos, *vel, *mass, *rho, *P, *cs, *H, *omega;
  const double *fCl, *fCq, *fsh;   // Q multipliers (anisotropic records), or null
};
// Where the records go: rec (FP64 record), sec (SEC bit 1: mixed, bit 2: FP64
// isotropic pair record), sec2 (SEC bit 4: FP64 density record), sec3 (SEC
// bit 8: mixed density record); n slots.
struct PackOut {
  double* rec;
  void* sec;
  void* sec2;
  int64_t n, nInt;
  void* sec3;
  double qInv;     // 1/u of the 3-D mixed record's positions
};

// The records of slot k.  SEC (bits): also write, from the values in
// registers, the mixed pair kernel's record (1), the FP64 isotropic pair
// kernel's record (2), the FP64 density sum's record (4) and the mixed density
// sum's record (8).
template <int D, bool ISO, int SEC>
__device__ __forceinline__ void packSlot(int64_t k, const PackIn& in, const PackOut& o) {
  static_assert(SEC == 0 || ISO, "the pair kernels' records are isotropic");
  using R = RL<ISO>;
  constexpr int NS = NSym<D>::value;
  const int64_t i = in.perm[k];
  double r[R::STRIDE];
#pragma unroll
  for (int c = 0; c < R::STRIDE; ++c) r[c] = 0.0;
  if (i < 0) {                      // dead slot (culled ghost): no row holds it
    r[R::ORIG] = 2147483648.0;
    double2* dst = reinterpret_cast<double2*>(o.rec + k * R::STRIDE);
#pragma unroll
    for (int c = 0; c < R::STRIDE / 2; ++c) dst[c] = make_double2(r[2 * c], r[2 * c + 1]);
    if constexpr ((SEC & 1) != 0) writeMixedRec<D>(r, o.nInt, o.qInv, static_cast<float4*>(o.sec) + k * MixedRec<D>::F4);
    if constexpr ((SEC & 2) != 0)
      writeFast64Rec<D>(r, o.nInt, static_cast<double2*>(o.sec) + k * 4, static_cast<double2*>(o.sec) + 4 * o.n + k * 2);
    if constexpr ((SEC & 4) != 0) writeDensRec<D>(r, static_cast<double2*>(o.sec2) + k * 3);
    if constexpr ((SEC & 8) != 0) writeDensRecMixed<D>(r, static_cast<float4*>(o.sec3) + k * DensMixedRec<D>::F4);
    return;
  }
  for (int c = 0; c < D; ++c) { r[R::X + c] = in.pos[i * D + c]; r[R::V + c] = in.vel[i * D + c]; }
  S<D> h;
  for (int c = 0; c < NS; ++c) h.c[c] = in.H[i * NS + c];
  const double rh = in.rho[i], m = in.mass[i], p = in.P[i];
  const double som = safeInv(in.omega[i], 1.0e-30);
  if constexpr (ISO) {
    r[RL<true>::HINV] = h.c[0];
  } else {
    for (int c = 0; c < NS; ++c) r[RL<false>::H + c] = h.c[c];
    r[RL<false>::FCL] = in.fCl ? in.fCl[i] : 1.0;
    r[RL<false>::FCQ] = in.fCq ? in.fCq[i] : 1.0;
    r[RL<false>::FSH] = in.fsh ? in.fsh[i] : 1.0;
  }
  r[R::HDET] = sdet(h);
  r[R::M] = m;
  r[R::ORIG] = double(i);
  r[R::RHO] = rh;
  r[R::CS] = in.cs[i];
  r[R::P] = p;
  r[R::SOM] = som;
  r[R::PRHO] = som * p / (rh * rh);
  r[R::MOR] = m / rh;
  double2* dst = reinterpret_cast<double2*>(o.rec + k * R::STRIDE);
#pragma unroll
  for (int c = 0; c < R::STRIDE / 2; ++c) dst[c] = make_double2(r[2 * c], r[2 * c + 1]);
  if constexpr ((SEC & 1) != 0) writeMixedRec<D>(r, o.nInt, o.qInv, static_cast<float4*>(o.sec) + k * MixedRec<D>::F4);
  if constexpr ((SEC & 2) != 0)
    writeFast64Rec<D>(r, o.nInt, static_cast<double2*>(o.sec) + k * 4, static_cast<double2*>(o.sec) + 4 * o.n + k * 2);
  if constexpr ((SEC & 4) != 0) writeDensRec<D>(r, static_cast<double2*>(o.sec2) + k * 3);
  if constexpr ((SEC & 8) != 0) writeDensRecMixed<D>(r, static_cast<float4*>(o.sec3) + k * DensMixedRec<D>::F4);
}

struct Sim;
PackIn packInputs(Sim& s);
// allocates the records that sec asks for; sets recSValid and recSMValid from bits 4 and 8
PackOut packOutputs(Sim& s, int sec);
// 1/u for the 3-D mixed record's positions (a power of two; from the last
// neighbour build's largest search radius)
double quantInv(const Sim& s);
// The records of the ghost slots only (after the ghost rows' fields changed).
void packGhostRecords(Sim& s, int sec);

}  // namespace sphgpu
