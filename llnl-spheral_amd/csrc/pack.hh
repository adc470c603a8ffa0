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
// The isotropic pair kernels' records, from an FP64 record r (isotropic
// layout): FP32, 64 B, for the mixed kernels (positions keep their FP64 bits);
// FP64, 96 B in two planes, for the FP64 kernel.  Ghost rows are marked by a
// negative det H (mixed) or 1/h (FP64).
template <int D>
__device__ __forceinline__ void writeMixedRec(const double* r, int64_t nInt, float4* __restrict__ p) {
  using R = RL<true>;
  const double x0 = r[R::X], x1 = D > 1 ? r[R::X + 1] : 0.0, x2 = D > 2 ? r[R::X + 2] : 0.0;
  const bool ghost = r[R::ORIG] >= double(nInt);
  const float hdet = float(r[R::HDET]);
  p[0] = make_float4(__int_as_float(__double2loint(x0)), __int_as_float(__double2hiint(x0)),
                     __int_as_float(__double2loint(x1)), __int_as_float(__double2hiint(x1)));
  p[1] = make_float4(__int_as_float(__double2loint(x2)), __int_as_float(__double2hiint(x2)), float(r[R::HINV]),
                     ghost ? -hdet : hdet);
  const double rho = r[R::RHO];
  p[2] = make_float4(float(r[R::M]), rho != 0.0 ? float(1.0 / rho) : 0.0f, float(r[R::V]),
                     D > 1 ? float(r[R::V + 1]) : 0.0f);
  p[3] = make_float4(D > 2 ? float(r[R::V + 2]) : 0.0f, float(r[R::CS]), float(r[R::PRHO]), float(r[R::MOR]));
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

// The mixed density sum's neighbour record, 32 B: x (FP64 bits), 1/h and m
// (FP32); det H = (1/h)^D is formed in FP32 by the reader.
template <int D>
__device__ __forceinline__ void writeDensRecMixed(const double* r, float4* __restrict__ p) {
  using R = RL<true>;
  const double x0 = r[R::X], x1 = D > 1 ? r[R::X + 1] : 0.0, x2 = D > 2 ? r[R::X + 2] : 0.0;
  p[0] = make_float4(__int_as_float(__double2loint(x0)), __int_as_float(__double2hiint(x0)),
                     __int_as_float(__double2loint(x1)), __int_as_float(__double2hiint(x1)));
  p[1] = make_float4(__int_as_float(__double2loint(x2)), __int_as_float(__double2hiint(x2)), float(r[R::HINV]),
                     float(r[R::M]));
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
    if constexpr ((SEC & 1) != 0) writeMixedRec<D>(r, o.nInt, static_cast<float4*>(o.sec) + k * 4);
    if constexpr ((SEC & 2) != 0)
      writeFast64Rec<D>(r, o.nInt, static_cast<double2*>(o.sec) + k * 4, static_cast<double2*>(o.sec) + 4 * o.n + k * 2);
    if constexpr ((SEC & 4) != 0) writeDensRec<D>(r, static_cast<double2*>(o.sec2) + k * 3);
    if constexpr ((SEC & 8) != 0) writeDensRecMixed<D>(r, static_cast<float4*>(o.sec3) + k * 2);
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
  if constexpr ((SEC & 1) != 0) writeMixedRec<D>(r, o.nInt, static_cast<float4*>(o.sec) + k * 4);
  if constexpr ((SEC & 2) != 0)
    writeFast64Rec<D>(r, o.nInt, static_cast<double2*>(o.sec) + k * 4, static_cast<double2*>(o.sec) + 4 * o.n + k * 2);
  if constexpr ((SEC & 4) != 0) writeDensRec<D>(r, static_cast<double2*>(o.sec2) + k * 3);
  if constexpr ((SEC & 8) != 0) writeDensRecMixed<D>(r, static_cast<float4*>(o.sec3) + k * 2);
}

struct Sim;
PackIn packInputs(Sim& s);
// allocates the records that sec asks for; sets recSValid and recSMValid from bits 4 and 8
PackOut packOutputs(Sim& s, int sec);
// The records of the ghost slots only (after the ghost rows' fields changed).
void packGhostRecords(Sim& s, int sec);

}  // namespace sphgpu
