// Time integration on the device: GammaLawGas, GenericHydro::dt and
// SynchronousRK2::step with the SPH state policies.
//
// Step (src/Integrator/SynchronousRK2.cc:78-144):
//   ghosts + neighbours (Integrator::preStepInitialize -> setGhostNodes),
//   dt = selectDt(GenericHydro::dt) with dtGrowth and the dtMin/dtMax clamps,
//   derivatives at t, state0 <- state, update by dt/2, ghosts,
//   derivatives at t + dt/2, state <- state0 updated by dt with them, ghosts,
//   SPHHydroBase::finalize (sum density with the step's neighbour rows),
//   Integrator::enforceBoundaries.
// State::update with the policies of SPHHydroBase::registerState
// (src/SPH/SPHHydroBase.cc:238-336) on the internal rows:
//   x += m*DxDt, v += m*DvDt, eps += m*DepsDt          (IncrementFieldList)
//   rho = min(rhoMax, max(rhoMin, rho + m*DrhoDt))      (IncrementBoundedState)
//   H = ideal H with eigenvalues bounded by 1/hmax, 1/hmin (ReplaceBoundedState;
//       the SPH ideal H is h1^-1 I, so the bound is a clamp of the diagonal)
//   P, c from the gamma-law EOS                         (Pressure/SoundSpeedPolicy)
// With compatibleEnergyEvolution, eps follows SpecificThermalEnergyPolicy
// (src/Hydro/SpecificThermalEnergyPolicy.cc:64-189): eps += m*DepsDt with
// DepsDt rebuilt from the pair accelerations of the pass and the step-start
// velocities.
// The derivative fields are read in slot order (one lane per internal slot of
// the work list) and the state is written at the caller index of the slot.
#include "internal.hh"
#include "dgeom.hh"
#include "eigen.hh"
#include "ell.hh"
#include "pack.hh"
#include <cfloat>
#include <cstring>
#include <algorithm>
#include <mutex>

namespace sphgpu {

namespace {

constexpr int kBlock = 256;
inline int nblocks(int64_t n, int bs = kBlock) { return int(std::max<int64_t>(1, (n + bs - 1) / bs)); }

struct Eos {
  double gamma, gamma1, Pmin, Pmax;
  int floor;
};
// GammaLawGas::pressure with applyPressureLimits, soundSpeed
__device__ __forceinline__ double eosP(const Eos& e, double rho, double eps) {
  const double P = e.gamma1 * rho * eps;
  return P < e.Pmin ? (e.floor ? e.Pmin : 0.0) : (P > e.Pmax ? e.Pmax : P);
}
__device__ __forceinline__ double eosCs(const Eos& e, double eps) { return sqrt(dmax(0.0, e.gamma * e.gamma1 * eps)); }

__global__ void k_eos(const double* __restrict__ rho, const double* __restrict__ eps, double* __restrict__ P,
                      double* __restrict__ cs, int64_t n, Eos e) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  P[i] = eosP(e, rho[i], eps[i]);
  cs[i] = eosCs(e, eps[i]);
}

struct Fields {
  double *pos, *vel, *eps, *rho, *H, *P, *cs;
};
struct Base {
  const double *pos, *vel, *eps, *rho;
};
struct DerivIn {
  const double* base;  // field-major blocks of the last evaluation, slot rows
  int64_t n;
  int oDx, oDv, oDeps, oDrho, oHid;
  const double* compDe;  // compatible energy: DepsDt of the policy per slot, else null
};
struct Bounds {
  double rhoMin, rhoMax, hmaxInv, hminInv;
};
// Ensemble members: the update's multiplier (dt or dt/2) and gamma per member.
struct SegArgs {
  const double* seg;        // member of each row; null: one simulation
  SegState* st;             // null: the scalar multiplier and EOS apply to every row
  int half;                 // 1: multiply by dt/2
  const double* listOf = nullptr;      // several NodeLists with their own gamma-law EOS:
  const double* listGamma = nullptr;   //   each row's NodeList, and the NodeLists' gammas
};
__device__ __forceinline__ void rowMultEos(const SegArgs& sa, int64_t i, double& mult, Eos& e) {
  if (sa.listGamma) {
    const double g = sa.listGamma[int(sa.listOf[i])];
    e.gamma = g;
    e.gamma1 = g - 1.0;
  }
  if (!sa.st) return;
  if (!sa.seg) {            // one simulation whose dt was chosen on the device
    mult = sa.half ? sa.st->hdt : sa.st->dt;
    return;
  }
  const SegState& g = sa.st[int(sa.seg[i])];
  mult = sa.half ? g.hdt : g.dt;
  e.gamma = g.gamma;
  e.gamma1 = g.gamma - 1.0;
}

// REC >= 0: also write the slot's records (pack sec bits REC, isotropic
// layout) from the values just stored, as the record pack would.
template <int D, int REC>
__global__ void k_update(const int32_t* __restrict__ work, const int32_t* __restrict__ perm, int64_t nInt, Fields f,
                         Base b, DerivIn d, double mult, Bounds bd, Eos e, SegArgs sa, PackIn pin,
                         PackOut pout) {
  constexpr int NS = NSym<D>::value;
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= nInt) return;
  const int64_t k = work[t];
  const int64_t i = perm[k];
  rowMultEos(sa, i, mult, e);
  const double* dx = d.base + int64_t(d.oDx) * d.n + k * D;
  const double* dv = d.base + int64_t(d.oDv) * d.n + k * D;
  for (int c = 0; c < D; ++c) {
    f.pos[i * D + c] = b.pos[i * D + c] + mult * dx[c];
    f.vel[i * D + c] = b.vel[i * D + c] + mult * dv[c];
  }
  const double eps = d.compDe ? b.eps[i] + d.compDe[k] * mult : b.eps[i] + mult * d.base[int64_t(d.oDeps) * d.n + k];
  const double rho = dmin(bd.rhoMax, dmax(bd.rhoMin, b.rho[i] + mult * d.base[int64_t(d.oDrho) * d.n + k]));
  f.eps[i] = eps;
  f.rho[i] = rho;
  const double* hid = d.base + int64_t(d.oHid) * d.n + k * NS;
  int c = 0;
  for (int a = 0; a < D; ++a)
    for (int q = a; q < D; ++q, ++c) f.H[i * NS + c] = (a == q) ? dmin(bd.hminInv, dmax(bd.hmaxInv, hid[c])) : hid[c];
  const double P = eosP(e, rho, eps), cs = eosCs(e, eps);
  f.P[i] = P;
  f.cs[i] = cs;
  if constexpr (REC >= 0) packSlot<D, true, REC>(k, pin, pout);
  // an ensemble member whose state turns non-finite fails alone (the step's
  // final update; NaN at the midpoint reaches it)
  if (sa.seg && !sa.half) {
    bool ok = isfinite(eps) && isfinite(rho) && isfinite(P) && isfinite(cs);
    for (int c = 0; c < D; ++c) ok = ok && isfinite(f.pos[i * D + c]) && isfinite(f.vel[i * D + c]);
    for (int c = 0; c < NS; ++c) ok = ok && isfinite(f.H[i * NS + c]);
    if (!ok) atomicOr(&sa.st[int(sa.seg[i])].fail, unsigned(SPHGPU_MEMBER_NONFINITE));
  }
}

// maxViscousPressure, tr DvDx and DvDt of the last evaluation, by caller index.
template <int D>
__global__ void k_dt_inputs(const int32_t* __restrict__ work, const int32_t* __restrict__ perm, int64_t nInt,
                            const double* __restrict__ out, int64_t n, int oVP, int oDvDx, int oDvDt,
                            double* __restrict__ dtIn) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= nInt) return;
  const int64_t k = work[t];
  const int64_t i = perm[k];
  constexpr int W = 2 + D;
  dtIn[i * W] = out[int64_t(oVP) * n + k];
  const double* g = out + int64_t(oDvDx) * n + k * D * D;
  double tr = g[0];
  for (int a = 1; a < D; ++a) tr = tr + g[a * D + a];
  dtIn[i * W + 1] = tr;
  for (int c = 0; c < D; ++c) dtIn[i * W + 2 + c] = out[int64_t(oDvDt) * n + k * D + c];
}

template <int D> __device__ __forceinline__ double vmagArr(const double* a) {
  if constexpr (D == 1) return dabs(a[0]);
  double s = a[0] * a[0];
  for (int c = 1; c < D; ++c) s = s + a[c] * a[c];
  return sqrt(s);
}

// GenericHydro::dt candidates of one internal node; the minimum over all
// nodes is reduced with one 64-bit atomic min per wavefront on the bit
// pattern of the (positive) candidate.
template <int D>
__global__ void k_dt(const int32_t* __restrict__ work, const int32_t* __restrict__ perm, int64_t nInt,
                     const int32_t* __restrict__ rowCount, const int32_t* __restrict__ nbr, int64_t n,
                     const double* __restrict__ H, const double* __restrict__ cs, const double* __restrict__ rho,
                     const double* __restrict__ vel, const double* __restrict__ dtIn, int haveDtIn, double nPerh,
                     int useVelMag, const double* __restrict__ seg, unsigned long long* __restrict__ minBits,
                     int64_t minStride) {
  constexpr int NS = NSym<D>::value;
  const int64_t t = xcdBlock((nInt + kBlock - 1) / kBlock) * kBlock + threadIdx.x;
  double m = double(FLT_MAX);
  int sg = 0;
  if (t < nInt) {
    const int64_t k = work[t];
    const int64_t i = perm[k];
    if (seg) sg = int(seg[i]);
    const double nodeScale = 1.0 / maxEigen<D>(H + i * NS) / nPerh;
    m = dmin(m, nodeScale / (cs[i] + FLT_MIN));
    constexpr int W = 2 + D;
    double vp = 0.0, divv = 0.0, acc[3] = {0.0, 0.0, 0.0};
    if (haveDtIn) {
      vp = dtIn[i * W];
      divv = dtIn[i * W + 1];
      for (int c = 0; c < D; ++c) acc[c] = dtIn[i * W + 2 + c];
    }
    const double csq = sqrt(vp / rho[i]);
    m = dmin(m, nodeScale / (csq + FLT_MIN));
    m = dmin(m, 1.0 / (dabs(divv) + FLT_MIN));
    double vi[3];
    for (int c = 0; c < D; ++c) vi[c] = vel[i * D + c];
    const int len = rowCount[k];
    for (int w = 0; w < len; ++w) {
      const int64_t j = perm[nbr[ellIndex(w, n, k)]];
      double d[3];
      for (int c = 0; c < D; ++c) d[c] = vi[c] - vel[j * D + c];
      m = dmin(m, nodeScale * (1.0 / dmax(1.0e-30, vmagArr<D>(d))));
    }
    m = dmin(m, sqrt(nodeScale / (vmagArr<D>(acc) + FLT_MIN)));
    if (useVelMag) m = dmin(m, nodeScale / (vmagArr<D>(vi) + 1.0e-10));
  }
  const int s0 = __shfl(sg, 0);
  if (__all(sg == s0)) {
    for (int sh = 32; sh > 0; sh >>= 1) m = dmin(m, __shfl_xor(m, sh));
    if ((threadIdx.x & 63) == 0) atomicMin(minBits + s0 * minStride, (unsigned long long)__double_as_longlong(m));
  } else if (t < nInt) {
    atomicMin(minBits + sg * minStride, (unsigned long long)__double_as_longlong(m));
  }
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }
__global__ void k_fill_value(double* __restrict__ a, int64_t n, double v) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) a[i] = v;
}

// ---- SpecificThermalEnergyPolicy ------------------------------------------
// DvDt and DepsDt of the internal slots to caller order; the ghost rows are
// then filled by the boundaries (SPHHydroBase::finalizeDerivatives).
template <int D>
__global__ void k_compat_scatter(const int32_t* __restrict__ work, const int32_t* __restrict__ perm, int64_t nInt,
                                 const double* __restrict__ out, int64_t n, int oDv, int oDeps,
                                 double* __restrict__ acc, double* __restrict__ de0) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= nInt) return;
  const int64_t k = work[t];
  const int64_t i = perm[k];
  for (int c = 0; c < D; ++c) acc[i * D + c] = out[int64_t(oDv) * n + k * D + c];
  de0[i] = out[int64_t(oDeps) * n + k];
}

// Per slot: v + a hdt, |DepsDt0| + epsilon, m and the caller index (the pair
// owner is the node of smaller caller index).  Dead slots are in no row.
template <int D>
__global__ void k_compat_rec(const int32_t* __restrict__ perm, int64_t nSlots, const double* __restrict__ vel0,
                             const double* __restrict__ acc, const double* __restrict__ de0,
                             const double* __restrict__ mass, double mult, SegArgs sa, double* __restrict__ crec) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k >= nSlots) return;
  const int64_t i = perm[k];
  double r[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i >= 0) {
    Eos unused{};
    rowMultEos(sa, i, mult, unused);
    const double hdt = 0.5 * mult;
    for (int c = 0; c < D; ++c) r[c] = vel0[i * D + c] + acc[i * D + c] * hdt;
    r[3] = dabs(de0[i]) + DBL_EPSILON;
    r[4] = mass[i];
    r[5] = double(i);
  }
  double2* o = reinterpret_cast<double2*>(crec + k * 8);
  o[0] = make_double2(r[0], r[1]);
  o[1] = make_double2(r[2], r[3]);
  o[2] = make_double2(r[4], r[5]);
}

// DepsDt of one internal node, gathered along its row.  The row holds the
// pass's own-frame pair term dv of each neighbour j; the reference's pair
// accelerations are -m_j dv for the node's own pairs (i owns when its caller
// index is the smaller) and m_i dv for the neighbour's end of a pair j owns.
// The owner end receives w_i du, the other end (1 - w_i) du m_i/m_j, written
// here from the receiving node with the owner's operands in the reference's
// order.  With rows in caller-index order the sum runs in the reference's
// order (owners j < i ascending, then the node's own pairs).
template <int D>
__global__ void k_compat(const int32_t* __restrict__ work, const int32_t* __restrict__ rowCount,
                         const int32_t* __restrict__ nbr, int64_t n, int64_t nInt, const double* __restrict__ crec,
                         const double* __restrict__ pairDv, int64_t plane, double* __restrict__ compDe) {
  const int64_t t = xcdBlock((nInt + kBlock - 1) / kBlock) * kBlock + threadIdx.x;
  if (t >= nInt) return;
  const int64_t k = work[t];
  const double2* mr = reinterpret_cast<const double2*>(crec + k * 8);
  const double2 a0 = mr[0], a1 = mr[1], a2 = mr[2];
  const double vi[3] = {a0.x, a0.y, a1.x};
  const double wti = a1.y, mi = a2.x, oi = a2.y;
  double de = 0.0;
  const int len = rowCount[k];
  for (int w = 0; w < len; ++w) {
    const int64_t e = ellIndex(w, n, k);
    const double2* nr = reinterpret_cast<const double2*>(crec + int64_t(nbr[e]) * 8);
    const double2 b0 = nr[0], b1 = nr[1], b2 = nr[2];
    const double vj[3] = {b0.x, b0.y, b1.x};
    const double wtj = b1.y, mj = b2.x, oj = b2.y;
    double du = 0.0;
    if (oj > oi) {
      for (int c = 0; c < D; ++c) {
        const double pa = pairDv[int64_t(c) * plane + e] * (-mj);
        const double tc = (vj[c] - vi[c]) * pa;
        du = (c == 0) ? tc : du + tc;
      }
      const double wi = wti / (wti + wtj);
      de += wi * du;
    } else {
      for (int c = 0; c < D; ++c) {
        const double pa = pairDv[int64_t(c) * plane + e] * mi;
        const double tc = (vi[c] - vj[c]) * pa;
        du = (c == 0) ? tc : du + tc;
      }
      const double wj = wtj / (wtj + wti);
      de += (1.0 - wj) * du * mj / mi;
    }
  }
  compDe[k] = de;
}

Eos makeEos(double gamma, double Pmin, double Pmax, int floor) { return Eos{gamma, gamma - 1.0, Pmin, Pmax, floor}; }

// SpecificThermalEnergyPolicy::update's DepsDt into s.compDe (slot rows) for
// an update by `mult`, with vel0 the velocities the state holds before it.
template <int D> void compatD(Sim& s, const double* vel0, double mult, const SegArgs& sa) {
  hipStream_t st = s.stream;
  const size_t nt = size_t(std::max<int64_t>(s.nTot, 1)), ns = size_t(std::max<int64_t>(s.nSlots, 1));
  s.compAcc.ensure(sizeof(double) * D * nt);
  s.compDe0.ensure(sizeof(double) * nt);
  s.compRec.ensure(sizeof(double) * 8 * ns);
  s.compDe.ensure(sizeof(double) * ns);
  hipLaunchKernelGGL(k_compat_scatter<D>, dim3(nblocks(s.nInt)), dim3(kBlock), 0, st, s.work.as<int32_t>(),
                     s.perm.as<int32_t>(), s.nInt, s.out.as<double>(), s.outStride,
                     derivOffsetDoubles(D, SPHGPU_D_DVDT), derivOffsetDoubles(D, SPHGPU_D_DEPSDT),
                     s.compAcc.as<double>(), s.compDe0.as<double>());
  applyGhostVector(s, s.compAcc);
  applyGhostScalar(s, s.compDe0);
  hipLaunchKernelGGL(k_compat_rec<D>, dim3(nblocks(s.nSlots)), dim3(kBlock), 0, st, s.perm.as<int32_t>(), s.nSlots,
                     vel0, s.compAcc.as<double>(), s.compDe0.as<double>(), s.mass.as<double>(), mult, sa,
                     s.compRec.as<double>());
  hipLaunchKernelGGL(k_compat<D>, dim3(xcdGrid(nblocks(s.nInt))), dim3(kBlock), 0, st, s.work.as<int32_t>(),
                     s.rowCount.as<int32_t>(), s.nbr.as<int32_t>(), s.nSlots, s.nInt, s.compRec.as<double>(),
                     s.pairDv.as<double>(), s.nSlots * int64_t(s.ellCap), s.compDe.as<double>());
  SG_CHECK_HIP(hipGetLastError());
}

// rec >= 0: the update also writes the internal slots' records (pack sec bits;
// isotropic H only) and the caller packs the ghost slots' records after the
// ghost rows are refreshed.
template <int D> void updateD(Sim& s, const Base& b, double mult, const Bounds& bd, const Eos& e,
                              const double* compDe, const SegArgs& sa, int rec = -1) {
  Fields f{s.pos.as<double>(), s.vel.as<double>(), s.eps.as<double>(), s.rho.as<double>(), s.H.as<double>(),
           s.P.as<double>(), s.cs.as<double>()};
  DerivIn d{s.out.as<double>(), s.outStride, derivOffsetDoubles(D, SPHGPU_D_DXDT), derivOffsetDoubles(D, SPHGPU_D_DVDT),
            derivOffsetDoubles(D, SPHGPU_D_DEPSDT), derivOffsetDoubles(D, SPHGPU_D_DRHODT),
            derivOffsetDoubles(D, SPHGPU_D_HIDEAL), compDe};
  SG_REQUIRE(rec < 0 || s.isotropic, SPHGPU_ERR_STATE, "records from the update need H = h I");
  const PackIn pin = rec >= 0 ? packInputs(s) : PackIn{};
  const PackOut pout = rec >= 0 ? packOutputs(s, rec) : PackOut{};
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nblocks(s.nInt)), dim3(kBlock), 0, s.stream, s.work.as<int32_t>(),
                       s.perm.as<int32_t>(), s.nInt, f, b, d, mult, bd, e, sa, pin, pout);
  };
  switch (rec) {
    case -1: go(k_update<D, -1>); break;
    case 0: go(k_update<D, 0>); break;
    case 1: go(k_update<D, 1>); break;
    case 2: go(k_update<D, 2>); break;
    case 4: go(k_update<D, 4>); break;
    case 8: go(k_update<D, 8>); break;
    default: SG_REQUIRE(false, SPHGPU_ERR_STATE, "unsupported record kinds for the update");
  }
  SG_CHECK_HIP(hipGetLastError());
  s.recValid = false;
}

template <int D> void dtInputsD(Sim& s) {
  s.dtIn.ensure(sizeof(double) * size_t(2 + D) * size_t(std::max<int64_t>(s.nInt, 1)));
  hipLaunchKernelGGL(k_dt_inputs<D>, dim3(nblocks(s.nInt)), dim3(kBlock), 0, s.stream, s.work.as<int32_t>(),
                     s.perm.as<int32_t>(), s.nInt, s.out.as<double>(), s.outStride,
                     derivOffsetDoubles(D, SPHGPU_D_MAX_VISCOUS_PRESSURE), derivOffsetDoubles(D, SPHGPU_D_DVDX),
                     derivOffsetDoubles(D, SPHGPU_D_DVDT), s.dtIn.as<double>());
  SG_CHECK_HIP(hipGetLastError());
  s.haveDtIn = true;
}

template <int D> double hydroDtD(Sim& s, double nPerh, double cfl, int useVelMag) {
  s.reduceTmp.ensure(sizeof(double) * 16);
  unsigned long long* bits = reinterpret_cast<unsigned long long*>(s.reduceTmp.as<double>() + 9);
  const double init = double(FLT_MAX);
  unsigned long long ib;
  std::memcpy(&ib, &init, sizeof(ib));
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s.stream, bits, ib);
  if (!s.haveDtIn) s.dtIn.ensure(sizeof(double) * 8);
  hipLaunchKernelGGL(k_dt<D>, dim3(xcdGrid(nblocks(s.nInt))), dim3(kBlock), 0, s.stream, s.work.as<int32_t>(),
                     s.perm.as<int32_t>(), s.nInt, s.rowCount.as<int32_t>(), s.nbr.as<int32_t>(), s.nSlots,
                     s.H.as<double>(), s.cs.as<double>(), s.rho.as<double>(), s.vel.as<double>(), s.dtIn.as<double>(),
                     s.haveDtIn ? 1 : 0, nPerh, useVelMag, nullptr, bits, int64_t(0));
  SG_CHECK_HIP(hipGetLastError());
  unsigned long long hb = 0;
  SG_CHECK_HIP(hipMemcpyAsync(&hb, bits, sizeof(hb), hipMemcpyDeviceToHost, s.stream));
  SG_CHECK_HIP(hipStreamSynchronize(s.stream));
  double minDt;
  std::memcpy(&minDt, &hb, sizeof(minDt));
  return cfl * minDt;
}

// Per-member dt selection, the same expressions as rk2D's host selection.
__global__ void k_seg_vote_init(SegState* __restrict__ st, int n, unsigned long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    st[i].voteBits = v;
    st[i].fail = 0u;
  }
}
// dtThreshold > 0: a member whose vote falls below it (a collapsing step) is
// flagged; its step still runs and it is evicted afterwards.
__global__ void k_seg_select_dt(SegState* __restrict__ st, int n, double cfl, double dtMinIn, double dtMaxIn,
                                double dtGrowth, double dtThreshold) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  SegState& g = st[i];
  const double dtMin = dmin(dtMinIn, g.tEnd - g.t), dtMax = dmin(dtMaxIn, g.tEnd - g.t);
  const double vote = cfl * __longlong_as_double((long long)g.voteBits);
  if (dtThreshold > 0.0 && !(vote >= dtThreshold)) g.fail |= unsigned(SPHGPU_MEMBER_DT_COLLAPSE);
  double dt = dtMax;
  if (vote > 0.0 && vote < dt) dt = vote;
  dt = dmin(dt, dtGrowth * g.lastDt);
  dt = dmin(dtMax, dmax(dtMin, dt));
  g.dt = dt;
  g.hdt = 0.5 * dt;
}
__global__ void k_seg_advance(SegState* __restrict__ st, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  SegState& g = st[i];
  g.t = g.t + g.dt;
  g.lastDt = g.dt;
  g.cycle = g.cycle + 1.0;
}

// Several device copies in one dispatch (the step's saved state of a small
// node set: five launches cost more than the bytes).
struct CopySet {
  double* d[5];
  const double* s[5];
  int64_t n[5];
  int m;
};
constexpr size_t kCopyManyMaxBytes = size_t(32) << 20;
__global__ void k_copy_many(CopySet c) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int q = 0; q < c.m; ++q)
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < c.n[q]; i += stride) c.d[q][i] = c.s[q][i];
}

// Integrator::preStepInitialize: ghosts and connectivity.  The build's record
// pack also writes the pair kernel's record, which the step reads first (grad-h
// recomputes omega, which the records hold, before that).
void preStep(Sim& s, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip) {
  const bool mixed = hp.precision == SPHGPU_PRECISION_MIXED;
  const bool fast64 = s.rowOrder == SPHGPU_ROW_ORDER_SEARCH && hp.epsTensile == 0.0 && !hp.compatibleEnergyEvolution;
  struct Hint {     // the hint holds for this build only, also when it throws
    Sim& s;
    ~Hint() { s.buildRecHint = -1; }
  } hint{s};
  s.buildRecHint = ip.gradhCorrection ? 0 : (mixed ? 1 : (fast64 ? 2 : 0));
  if (!s.subs.empty()) setGhostNodes(s, ip.kernelExtent, ip.cullGhostNodes != 0);
  else buildNeighbors(s, ip.kernelExtent);
}

// The two derivative passes and updates of SynchronousRK2::step, then
// SPHHydroBase::finalize and Integrator::enforceBoundaries.  Multipliers are
// dt/2, dt (plain simulation) or per member (half, full).
template <int D> void stepBody(Sim& s, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip, double dt,
                               const SegArgs& half, const SegArgs& full) {
  constexpr int NS = NSym<D>::value;
  hipStream_t st = s.stream;
  const double hdt = 0.5 * dt;
  const bool compat = hp.compatibleEnergyEvolution != 0;
  SG_REQUIRE(!compat || s.deviceGhosts || s.nTot == s.nInt, SPHGPU_ERR_UNSUPPORTED,
             "compatible energy with caller-supplied ghost rows: the ghost derivatives need the device boundaries");
  const Bounds bd{ip.rhoMin, ip.rhoMax, 1.0 / hp.hmax, 1.0 / hp.hmin};
  const Eos e = makeEos(ip.gamma, ip.Pmin, ip.Pmax, ip.pressureFloor);

  // derivatives at t (SPHHydroBase::initialize first: the grad-h correction); the
  // midpoint pass below rewrites every field before anything but the trial
  // update reads one, so this pass stores only the fields the update reads
  if (ip.gradhCorrection) omegaGradh(s);
  evalDerivs(s, hp, true);
  // state0 (the policy reads every row's step-start velocity, ghosts included)
  const size_t n = size_t(s.nInt), nv = compat ? size_t(s.nTot) : n;
  const size_t bytes[5] = {sizeof(double) * n * D, sizeof(double) * nv * D, sizeof(double) * n, sizeof(double) * n,
                           sizeof(double) * n * NS};
  DBuf* src[5] = {&s.pos, &s.vel, &s.eps, &s.rho, &s.H};
  CopySet cs{};
  size_t total = 0;
  for (int q = 0; q < 5; ++q) {
    s.state0[q].ensure(std::max<size_t>(bytes[q], 8));
    cs.d[q] = s.state0[q].as<double>();
    cs.s[q] = src[q]->as<double>();
    cs.n[q] = int64_t(bytes[q] / sizeof(double));
    total += bytes[q];
  }
  cs.m = 5;
  if (total <= kCopyManyMaxBytes) {
    // one dispatch for the five copies of a small node set
    hipLaunchKernelGGL(k_copy_many, dim3(nblocks(std::max<int64_t>(1, int64_t(total / sizeof(double) / 5)))),
                       dim3(kBlock), 0, st, cs);
    SG_CHECK_HIP(hipGetLastError());
  } else {
    for (int q = 0; q < 5; ++q)
      if (bytes[q]) SG_CHECK_HIP(hipMemcpyAsync(s.state0[q].p, src[q]->p, bytes[q], hipMemcpyDeviceToDevice, st));
  }
  const Base b0{s.state0[0].as<double>(), s.state0[1].as<double>(), s.state0[2].as<double>(), s.state0[3].as<double>()};
  // omega has no update policy (SPHHydroBase::registerState enrolls it bare):
  // state.assign(state0) gives the final update the step-start value again
  const size_t omegaBytes = sizeof(double) * size_t(s.nTot);
  if (ip.gradhCorrection && omegaBytes) {
    s.omega0.ensure(omegaBytes);
    SG_CHECK_HIP(hipMemcpyAsync(s.omega0.p, s.omega.p, omegaBytes, hipMemcpyDeviceToDevice, st));
  }
  // The updates write the records the next pass reads (isotropic H; grad-h
  // recomputes omega, which the records hold, before the midpoint pass).
  const bool recs = s.isotropic && !ip.gradhCorrection;
  const int recMid = recs ? pairRecordSec(s, hp) : -1;
  const int recEnd = recs ? densityRecordSec(s) : -1;
  auto ghostsAndRecords = [&](int rec) {
    applyGhostBoundaries(s);
    if (rec >= 0) {
      packGhostRecords(s, rec);
      s.recValid = true;
    }
  };
  // trial advance to the midpoint
  if (compat) compatD<D>(s, b0.vel, hdt, half);
  updateD<D>(s, b0, hdt, bd, e, compat ? s.compDe.as<double>() : nullptr, half, recMid);
  ghostsAndRecords(recMid);
  // derivatives at t + dt/2 with the step's neighbour rows
  if (ip.gradhCorrection) omegaGradh(s);
  evalDerivs(s, hp);
  if (ip.gradhCorrection && omegaBytes)
    SG_CHECK_HIP(hipMemcpyAsync(s.omega.p, s.omega0.p, omegaBytes, hipMemcpyDeviceToDevice, st));
  // advance from state0 with the midpoint derivatives
  if (compat) compatD<D>(s, b0.vel, dt, full);
  updateD<D>(s, b0, dt, bd, e, compat ? s.compDe.as<double>() : nullptr, full, recEnd);
  ghostsAndRecords(recEnd);
  dtInputsD<D>(s);
  // SPHHydroBase::finalize: sum density
  sumDensity(s);
  // Integrator::enforceBoundaries
  enforceBoundaries(s);
}

// The clock of a plain simulation as one SegState, so that selectDt runs on
// the device (k_seg_select_dt) and the step needs no host round trip before
// its end.
__global__ void k_plain_seg_init(SegState* __restrict__ g, double t, double lastDt, double tEnd, double gamma,
                                 unsigned long long vote) {
  g->t = t;
  g->lastDt = lastDt;
  g->tEnd = tEnd;
  g->gamma = gamma;
  g->dt = 0.0;
  g->hdt = 0.0;
  g->voteBits = vote;
  g->cycle = 0.0;
  g->fail = 0u;
  g->failPad = 0u;
}

// ---- the device part of a step as a HIP graph ---------------------------------
// Everything between the neighbour build and the clock read-back (dt vote,
// dt selection, both passes, updates, ghost refreshes, finalize, boundary
// enforcement, clock advance) is launched without a host round trip.  Its
// launches depend on the node and slot counts, the device buffers, the
// parameters and the host-side record flags at its start: that is the key.
// The first step with a new key runs eagerly (sizing every buffer), the next
// step with the same key is captured and instantiated, and later steps with
// that key replay the graph; the host flags the step leaves are saved at the
// capture and restored after each replay.  An ensemble whose membership and
// ghost counts hold still (most steps of a tube ensemble) replays one graph
// per step.  Graphs hold no memset node (the pair count is cleared by a
// kernel) and no stage-timing events.
void keyBuf(std::vector<uint64_t>& k, const DBuf& b) {
  k.push_back(reinterpret_cast<uint64_t>(b.p));
  k.push_back(uint64_t(b.bytes));
}
template <class T> void keyBytes(std::vector<uint64_t>& k, const T& v) {
  const unsigned char* c = reinterpret_cast<const unsigned char*>(&v);
  for (size_t i = 0; i < sizeof(T); i += 8) {
    uint64_t w = 0;
    std::memcpy(&w, c + i, std::min<size_t>(8, sizeof(T) - i));
    k.push_back(w);
  }
}
std::vector<uint64_t> stepKey(const Sim& s, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip) {
  std::vector<uint64_t> k;
  k.reserve(256);
  for (int64_t v : {int64_t(s.dim), s.nInt, s.nTot, s.nSlots, int64_t(s.ellCap), int64_t(s.nSeg), s.tableVersion,
                    int64_t(s.isotropic), int64_t(s.rowOrder), int64_t(s.precision), int64_t(s.deviceGhosts),
                    int64_t(s.hasMultipliers), int64_t(s.recValid), int64_t(s.recSValid), int64_t(s.recSMValid),
                    int64_t(s.pairRec), int64_t(s.haveDtIn), int64_t(s.derivsValid), int64_t(s.subs.size()),
                    int64_t(s.nLists), int64_t(s.sumAllLists), int64_t(s.listGammaHost.size())})
    k.push_back(uint64_t(v));
  for (int64_t v : s.subStart) k.push_back(uint64_t(v));
  for (const SubBoundary& b : s.subs) keyBytes(k, b);
  keyBytes(k, hp);
  keyBytes(k, ip);
  keyBytes(k, s.dtThreshold);
  k.push_back(reinterpret_cast<uint64_t>(s.ncalcHost));
  for (const DBuf* b : {&s.pos, &s.vel, &s.mass, &s.rho, &s.eps, &s.P, &s.cs, &s.H, &s.omega, &s.fCl, &s.fCq,
                        &s.fshear, &s.cellOf, &s.cellCount, &s.cellStart, &s.cellFill, &s.perm, &s.iperm, &s.xs,
                        &s.Hs, &s.rowCount, &s.nbr, &s.scanTmp, &s.reduceTmp, &s.flags, &s.rec, &s.recM, &s.recD,
                        &s.recS, &s.recSM, &s.pairDv, &s.compDe, &s.compAcc, &s.compDe0, &s.compRec, &s.out,
                        &s.ghostControl, &s.ghostSub, &s.keep, &s.keepScan, &s.slotKeep, &s.hmaxOf, &s.genDev,
                        &s.work, &s.workFlag, &s.workScan, &s.omega0, &s.dtIn, &s.seg, &s.segState, &s.plainSeg,
                        &s.dTable, &s.dTable32, &s.dNperh, &s.dWsum, &s.pick, &s.listGamma})
    keyBuf(k, *b);
  for (const DBuf& b : s.state0) keyBuf(k, b);
  return k;
}

Sim::StepFlags stepFlagsOf(const Sim& s) {
  return Sim::StepFlags{s.recValid, s.recSValid, s.recSMValid, s.derivsValid, s.pairsValid, s.haveDtIn,
                        s.pairRec, s.outStride, s.timingPending};
}
void restoreStepFlags(Sim& s, const Sim::StepFlags& f) {
  s.recValid = f.recValid;
  s.recSValid = f.recSValid;
  s.recSMValid = f.recSMValid;
  s.derivsValid = f.derivsValid;
  s.pairsValid = f.pairsValid;
  s.haveDtIn = f.haveDtIn;
  s.pairRec = f.pairRec;
  s.outStride = f.outStride;
  s.timingPending = f.timingPending;
}

// A few graphs and recent keys are kept: culling swaps the state fields with
// their spares, so a step's buffers alternate between two sets.
constexpr size_t kStepGraphs = 4;

template <class F> void runStep(Sim& s, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip, F&& launches) {
  if (!s.useGraphs) {
    launches();
    return;
  }
  std::vector<uint64_t> key = stepKey(s, hp, ip);
  for (Sim::StepGraph& g : s.stepGraphs)
    if (g.key == key) {
      SG_CHECK_HIP(hipGraphLaunch(g.exec, s.stream));
      restoreStepFlags(s, g.flags);
      ++s.graphReplays;
      return;
    }
  if (std::find(s.recentKeys.begin(), s.recentKeys.end(), key) == s.recentKeys.end()) {
    // a new key: run eagerly, which also sizes every buffer the step uses
    launches();
    s.recentKeys.push_back(std::move(key));
    if (s.recentKeys.size() > kStepGraphs) s.recentKeys.erase(s.recentKeys.begin());
    return;
  }
  SG_CHECK_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
  s.capturing = true;
  try {
    launches();
  } catch (...) {
    s.capturing = false;
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  s.capturing = false;
  hipGraph_t g = nullptr;
  SG_CHECK_HIP(hipStreamEndCapture(s.stream, &g));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  SG_CHECK_HIP(e);
  if (s.stepGraphs.size() >= kStepGraphs) {
    SG_CHECK_HIP(hipGraphExecDestroy(s.stepGraphs.front().exec));
    s.stepGraphs.erase(s.stepGraphs.begin());
  }
  Sim::StepGraph sg;
  sg.key = std::move(key);
  sg.exec = exec;
  sg.flags = stepFlagsOf(s);
  s.stepGraphs.push_back(std::move(sg));
  SG_CHECK_HIP(hipGraphLaunch(exec, s.stream));
  ++s.graphCaptures;
}

template <int D> double rk2D(Sim& s, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip,
                             double maxTime) {
  hipStream_t st = s.stream;
  preStep(s, hp, ip);
  // GenericHydro::dt, then selectDt (dtMin/dtMax clamped to maxTime - t, dtGrowth)
  s.plainSeg.ensure(sizeof(SegState));
  SegState* sd = s.plainSeg.as<SegState>();
  const double init = double(FLT_MAX);
  unsigned long long ib;
  std::memcpy(&ib, &init, sizeof(ib));
  hipLaunchKernelGGL(k_plain_seg_init, dim3(1), dim3(1), 0, st, sd, s.time, s.lastDt, maxTime, ip.gamma, ib);
  if (!s.haveDtIn) s.dtIn.ensure(sizeof(double) * 8);
  runStep(s, hp, ip, [&] {
    hipLaunchKernelGGL(k_dt<D>, dim3(xcdGrid(nblocks(s.nInt))), dim3(kBlock), 0, st, s.work.as<int32_t>(),
                       s.perm.as<int32_t>(), s.nInt, s.rowCount.as<int32_t>(), s.nbr.as<int32_t>(), s.nSlots,
                       s.H.as<double>(), s.cs.as<double>(), s.rho.as<double>(), s.vel.as<double>(),
                       s.dtIn.as<double>(), s.haveDtIn ? 1 : 0, hp.nPerh, ip.useVelocityMagnitudeForDt, nullptr,
                       &sd->voteBits, int64_t(0));
    hipLaunchKernelGGL(k_seg_select_dt, dim3(1), dim3(kBlock), 0, st, sd, 1, ip.cfl, ip.dtMin, ip.dtMax, ip.dtGrowth,
                       0.0);
    SG_CHECK_HIP(hipGetLastError());
    const bool perList = s.nLists > 1 && !s.listGammaHost.empty();
    const double* lo = perList ? s.seg.as<double>() : nullptr;
    const double* lg = perList ? s.listGamma.as<double>() : nullptr;
    const SegArgs half{nullptr, sd, 1, lo, lg}, full{nullptr, sd, 0, lo, lg};
    stepBody<D>(s, hp, ip, 0.0, half, full);
    hipLaunchKernelGGL(k_seg_advance, dim3(1), dim3(kBlock), 0, st, sd, 1);
    SG_CHECK_HIP(hipGetLastError());
  });
  SegState* hp2 = static_cast<SegState*>(hostStaging(s, 0, sizeof(SegState)));
  SG_CHECK_HIP(hipMemcpyAsync(hp2, sd, sizeof(SegState), hipMemcpyDeviceToHost, st));
  SG_CHECK_HIP(hipStreamSynchronize(st));
  const SegState h = *hp2;
  s.time = h.t;
  s.lastDt = h.lastDt;
  s.cycle += 1;
  s.derivsValid = false;
  return h.dt;
}

// One SynchronousRK2 step of every ensemble member, each with its own dt.
template <int D> void ensembleRk2D(Sim& s, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip) {
  hipStream_t st = s.stream;
  const int ns = s.nSeg;
  SegState* sd = s.segState.as<SegState>();
  preStep(s, hp, ip);
  // GenericHydro::dt per member, then selectDt per member
  const double init = double(FLT_MAX);
  unsigned long long ib;
  std::memcpy(&ib, &init, sizeof(ib));
  if (!s.haveDtIn) s.dtIn.ensure(sizeof(double) * 8);
  runStep(s, hp, ip, [&] {
    hipLaunchKernelGGL(k_seg_vote_init, dim3(nblocks(ns)), dim3(kBlock), 0, st, sd, ns, ib);
    hipLaunchKernelGGL(k_dt<D>, dim3(xcdGrid(nblocks(s.nInt))), dim3(kBlock), 0, st, s.work.as<int32_t>(),
                       s.perm.as<int32_t>(), s.nInt, s.rowCount.as<int32_t>(), s.nbr.as<int32_t>(), s.nSlots,
                       s.H.as<double>(), s.cs.as<double>(), s.rho.as<double>(), s.vel.as<double>(),
                       s.dtIn.as<double>(), s.haveDtIn ? 1 : 0, hp.nPerh, ip.useVelocityMagnitudeForDt,
                       s.seg.as<double>(), &sd->voteBits, int64_t(sizeof(SegState) / sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_seg_select_dt, dim3(nblocks(ns)), dim3(kBlock), 0, st, sd, ns, ip.cfl, ip.dtMin, ip.dtMax,
                       ip.dtGrowth, s.dtThreshold);
    SG_CHECK_HIP(hipGetLastError());
    const SegArgs half{s.seg.as<double>(), sd, 1}, full{s.seg.as<double>(), sd, 0};
    stepBody<D>(s, hp, ip, 0.0, half, full);
    hipLaunchKernelGGL(k_seg_advance, dim3(nblocks(ns)), dim3(kBlock), 0, st, sd, ns);
    SG_CHECK_HIP(hipGetLastError());
  });
  SegState* hs = static_cast<SegState*>(hostStaging(s, 0, sizeof(SegState) * size_t(ns)));
  SG_CHECK_HIP(hipMemcpyAsync(hs, sd, sizeof(SegState) * size_t(ns), hipMemcpyDeviceToHost, st));
  SG_CHECK_HIP(hipStreamSynchronize(st));
  s.segHost.assign(hs, hs + ns);
  s.derivsValid = false;
}

__global__ void k_eos_rows(const double* __restrict__ rho, const double* __restrict__ eps, double* __restrict__ P,
                           double* __restrict__ cs, int64_t first, int64_t rows, Eos e) {
  const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (r >= rows) return;
  const int64_t i = first + r;
  P[i] = eosP(e, rho[i], eps[i]);
  cs[i] = eosCs(e, eps[i]);
}

}  // namespace

void eosGammaLaw(Sim& s, double gamma, double Pmin, double Pmax, int floorType) {
  if (s.nInt == 0) return;
  if (s.nLists > 1 && !s.listGammaHost.empty()) {
    // each NodeList with its own gamma-law gas
    for (int L = 0; L < s.nLists; ++L) {
      const int64_t first = s.listStart[size_t(L)], rows = s.listStart[size_t(L) + 1] - first;
      if (rows > 0) eosRows(s, first, rows, s.listGammaHost[size_t(L)], Pmin, Pmax, floorType);
    }
    SG_CHECK_HIP(hipStreamSynchronize(s.stream));
    s.recValid = false;
    s.derivsValid = false;
    return;
  }
  hipLaunchKernelGGL(k_eos, dim3(nblocks(s.nInt)), dim3(kBlock), 0, s.stream, s.rho.as<double>(), s.eps.as<double>(),
                     s.P.as<double>(), s.cs.as<double>(), s.nInt, makeEos(gamma, Pmin, Pmax, floorType));
  SG_CHECK_HIP(hipGetLastError());
  SG_CHECK_HIP(hipStreamSynchronize(s.stream));
  s.recValid = false;
  s.derivsValid = false;
}

double hydroDt(Sim& s, double nPerh, double cfl, int useVelMag) {
  SG_REQUIRE(s.neighborsValid, SPHGPU_ERR_STATE, "the dt vote needs a neighbour build");
  SG_REQUIRE(nPerh > 0.0, SPHGPU_ERR_ARG, "nPerh must be positive");
  if (s.dim == 1) return hydroDtD<1>(s, nPerh, cfl, useVelMag);
  if (s.dim == 2) return hydroDtD<2>(s, nPerh, cfl, useVelMag);
  return hydroDtD<3>(s, nPerh, cfl, useVelMag);
}

void ensembleRk2(Sim& s, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip) {
  if (s.dim == 1) ensembleRk2D<1>(s, hp, ip);
  else if (s.dim == 2) ensembleRk2D<2>(s, hp, ip);
  else ensembleRk2D<3>(s, hp, ip);
}

// ---- SpheralConservation totals ---------------------------------------------
// Per block, in a fixed order, over rows [r0, r1): mass, linear momentum (3),
// angular momentum about the origin (3), 2 x kinetic energy, thermal energy.
constexpr int kCons = 9;
template <int D>
__global__ void k_conserve_partial(const double* __restrict__ pos, const double* __restrict__ vel,
                                   const double* __restrict__ mass, const double* __restrict__ eps, int64_t r0,
                                   int64_t r1, double* __restrict__ partial) {
  __shared__ double sh[kCons][kBlock];
  double a[kCons] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = r0 + blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < r1; i += int64_t(gridDim.x) * blockDim.x) {
    const double m = mass[i];
    double x[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    for (int c = 0; c < D; ++c) { x[c] = pos[i * D + c]; v[c] = vel[i * D + c]; }
    a[0] += m;
    for (int c = 0; c < D; ++c) a[1 + c] += m * v[c];
    if (D == 2) a[6] += m * (x[0] * v[1] - x[1] * v[0]);
    if (D == 3) {
      a[4] += m * (x[1] * v[2] - x[2] * v[1]);
      a[5] += m * (x[2] * v[0] - x[0] * v[2]);
      a[6] += m * (x[0] * v[1] - x[1] * v[0]);
    }
    double v2 = v[0] * v[0];
    for (int c = 1; c < D; ++c) v2 = v2 + v[c] * v[c];
    a[7] += m * v2;
    a[8] += m * eps[i];
  }
  for (int q = 0; q < kCons; ++q) sh[q][threadIdx.x] = a[q];
  __syncthreads();
  for (int st = kBlock / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < kCons; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < kCons) partial[blockIdx.x * kCons + threadIdx.x] = sh[threadIdx.x][0];
}
__global__ void k_conserve_final(const double* __restrict__ partial, int nb, double* __restrict__ out) {
  const int q = threadIdx.x;
  if (q >= kCons) return;
  double t = 0.0;
  for (int b = 0; b < nb; ++b) t += partial[b * kCons + q];
  out[q] = t;
}

// The totals as reported: the kinetic energy is half the summed m v^2.
__global__ void k_conserve_report(const double* __restrict__ partial, int nb, double* __restrict__ out) {
  const int q = threadIdx.x;
  if (q >= kCons) return;
  double t = 0.0;
  for (int b = 0; b < nb; ++b) t += partial[b * kCons + q];
  out[q] = q == 7 ? 0.5 * t : t;
}

// Workgroups of the partial sums (the same for every row count, so the
// summation order depends only on the rows).
constexpr int kConserveBlocks = 256;
inline int conserveBlocks(int64_t rows) {
  return int(std::min<int64_t>(kConserveBlocks, std::max<int64_t>(1, (rows + kBlock - 1) / kBlock)));
}
void conservePartials(Sim& s, int64_t r0, int64_t r1, int nb, double* partial) {
  if (s.dim == 1)
    hipLaunchKernelGGL(k_conserve_partial<1>, dim3(nb), dim3(kBlock), 0, s.stream, s.pos.as<double>(),
                       s.vel.as<double>(), s.mass.as<double>(), s.eps.as<double>(), r0, r1, partial);
  else if (s.dim == 2)
    hipLaunchKernelGGL(k_conserve_partial<2>, dim3(nb), dim3(kBlock), 0, s.stream, s.pos.as<double>(),
                       s.vel.as<double>(), s.mass.as<double>(), s.eps.as<double>(), r0, r1, partial);
  else
    hipLaunchKernelGGL(k_conserve_partial<3>, dim3(nb), dim3(kBlock), 0, s.stream, s.pos.as<double>(),
                       s.vel.as<double>(), s.mass.as<double>(), s.eps.as<double>(), r0, r1, partial);
}

// Totals of rows [r0, r1) (internal nodes of the sim, or of one ensemble
// member): [mass, p (3), L (3), kinetic, thermal], as SpheralConservation
// (src/SimulationControl/SpheralConservation.py:66-150) takes them.
void conservationTotals(Sim& s, int64_t r0, int64_t r1, double* out) {
  const int nb = conserveBlocks(r1 - r0);
  DBuf tmp;
  tmp.ensure(sizeof(double) * size_t(nb + 1) * kCons);
  double* partial = tmp.as<double>();
  double* res = partial + size_t(nb) * kCons;
  conservePartials(s, r0, r1, nb, partial);
  hipLaunchKernelGGL(k_conserve_final, dim3(1), dim3(64), 0, s.stream, partial, nb, res);
  SG_CHECK_HIP(hipGetLastError());
  double h[kCons];
  SG_CHECK_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, s.stream));
  SG_CHECK_HIP(hipStreamSynchronize(s.stream));
  for (int q = 0; q < kCons; ++q) out[q] = h[q];
  out[7] = 0.5 * h[7];
}

// ---- the allocation-free totals ---------------------------------------------
// Partial sums of sphgpu_conservation_noalloc live in a file-scope device
// array instead of a per-call allocation: kConserveBlocks x kCons doubles
// (18 KB), held by every GPU the library runs on until the process ends.
// Each device has its own copy at its own address, resolved once per device
// with hipGetSymbolAddress (host code never passes the array's name itself).
// The partial kernel writes every slot it is read back from, so nothing needs
// clearing; calls that overlap on one GPU share the array and race.
constexpr size_t kScratchDoubles = size_t(kConserveBlocks) * kCons;
static __device__ __align__(256) double scratch[kScratchDoubles];
namespace {
constexpr int kMaxDevices = 64;
void* gScratch[kMaxDevices] = {};   // the array's address on each device
std::mutex gScratchMu;
}  // namespace

void resolveScratch(int device) {
  SG_REQUIRE(device >= 0 && device < kMaxDevices, SPHGPU_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(gScratchMu);
  if (gScratch[device]) return;
  void* p = nullptr;
  SG_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));   // the device is current
  gScratch[device] = p;
}

void conservationTotalsNoAlloc(Sim& s, int64_t r0, int64_t r1, double* outDevice) {
  void* p = nullptr;
  {
    std::lock_guard<std::mutex> lock(gScratchMu);
    p = (s.device >= 0 && s.device < kMaxDevices) ? gScratch[s.device] : nullptr;
  }
  SG_REQUIRE(p, SPHGPU_ERR_STATE, "the device's scratch array was not resolved");
  const int nb = conserveBlocks(r1 - r0);
  SG_REQUIRE(size_t(nb) * kCons <= kScratchDoubles, SPHGPU_ERR_CAPACITY,
             "conservation partial sums exceed the scratch array");
  double* partial = static_cast<double*>(p);
  conservePartials(s, r0, r1, nb, partial);
  hipLaunchKernelGGL(k_conserve_report, dim3(1), dim3(64), 0, s.stream, partial, nb, outDevice);
  SG_CHECK_HIP(hipGetLastError());
}

// iterateIdealH (src/Utilities/iterateIdealH.cc:30-310), SPH smoothing scale,
// nPerhForIteration = 0, no spherical start, no fixed determinant.  Each
// iteration makes the ghosts of the sim's boundaries (or only the
// connectivity without boundaries) from the current H and passes over the
// internal nodes; the loop ends after maxIterations or once the largest
// change is within tol.  The ghosts are remade from the final H at the end.
int iterateIdealH(Sim& s, const sphgpu_hydro_params& hp, int maxIterations, double tol, double ext, bool cull,
                  double* maxDelta) {
  SG_REQUIRE(s.nInt > 0, SPHGPU_ERR_STATE, "no nodes");
  SG_REQUIRE(!s.ensemble, SPHGPU_ERR_STATE, "ensembles iterate H per member alone");
  SG_REQUIRE(s.nTot == s.nInt || s.deviceGhosts, SPHGPU_ERR_STATE,
             "iterateIdealH makes its own ghosts: set_nodes(n_internal, n_internal)");
  SG_REQUIRE(maxIterations >= 0 && tol >= 0.0 && ext > 0.0, SPHGPU_ERR_ARG, "bad iteration parameters");
  DBuf deltaH, bits;
  deltaH.ensure(sizeof(double) * size_t(s.nInt));
  bits.ensure(sizeof(unsigned long long));
  const double init = 2.0 * tol;
  hipLaunchKernelGGL(k_fill_value, dim3(nblocks(s.nInt)), dim3(kBlock), 0, s.stream, deltaH.as<double>(), s.nInt, init);
  double m = 2.0 * tol;
  int it = 0;
  // several NodeLists: the pass walks each node's full connectivity, in
  // ConnectivityMap order (reference row order)
  struct MapOrder {
    Sim& s;
    explicit MapOrder(Sim& x) : s(x) { s.rowsMapOrder = s.nLists > 1; }
    ~MapOrder() { s.rowsMapOrder = false; }
  };
  {
    MapOrder mo(s);
    while (it < maxIterations && m > tol) {
      ++it;
      if (!s.subs.empty()) setGhostNodes(s, ext, cull);
      else buildNeighbors(s, ext);
      hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s.stream, bits.as<unsigned long long>(), 0ull);
      m = idealHPass(s, hp, tol, deltaH.as<double>(), bits.as<unsigned long long>());
    }
  }
  // leave the ghosts made from the final H (rows in the pass's order)
  if (!s.subs.empty()) setGhostNodes(s, ext, cull);
  else s.neighborsValid = false;
  SG_CHECK_HIP(hipStreamSynchronize(s.stream));
  if (maxDelta) *maxDelta = m;
  return it;
}

void eosRows(Sim& s, int64_t first, int64_t rows, double gamma, double Pmin, double Pmax, int floorType) {
  if (rows <= 0) return;
  hipLaunchKernelGGL(k_eos_rows, dim3(nblocks(rows)), dim3(kBlock), 0, s.stream, s.rho.as<double>(),
                     s.eps.as<double>(), s.P.as<double>(), s.cs.as<double>(), first, rows,
                     makeEos(gamma, Pmin, Pmax, floorType));
  SG_CHECK_HIP(hipGetLastError());
}

double rk2Step(Sim& s, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip, double maxTime) {
  SG_REQUIRE(!s.ensemble, SPHGPU_ERR_STATE, "an ensemble steps with sphgpu_ensemble_step");
  SG_REQUIRE(s.nInt > 0, SPHGPU_ERR_STATE, "no nodes");
  SG_REQUIRE(s.table.n > 0, SPHGPU_ERR_STATE, "no kernel table installed");
  SG_REQUIRE(ip.kernelExtent > 0.0 && ip.cfl > 0.0 && ip.dtGrowth > 0.0, SPHGPU_ERR_ARG, "bad integrator parameters");
  SG_REQUIRE(s.nTot == s.nInt || s.deviceGhosts, SPHGPU_ERR_STATE,
             "the integrator makes its own ghosts: set_nodes(n_internal, n_internal)");
  if (s.dim == 1) return rk2D<1>(s, hp, ip, maxTime);
  if (s.dim == 2) return rk2D<2>(s, hp, ip, maxTime);
  return rk2D<3>(s, hp, ip, maxTime);
}

}  // namespace sphgpu
