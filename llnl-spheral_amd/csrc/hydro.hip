// Device sum density and evaluateDerivatives.
//
// Both passes run in gather form: one lane per internal node (nodes in the
// neighbour build's sorted order), walking its full row.  The reference
// (src/SPH/SPHHydroBase.cc:575-928, src/SPH/computeSPHSumMassDensity.cc:60-103)
// evaluates each unordered pair once in the frame of its owner (the lower index,
// ConnectivityMapInline.hh:182-184) and adds to both ends, visiting owners in
// ascending order; so every node receives its pair terms in ascending neighbour
// order.  With rows in that order (SPHGPU_ROW_ORDER_REFERENCE), walking a row in
// the node's own frame reproduces the reference's FP64 sums exactly:
//   * quantities odd in r_ij (eta, unit vectors, grad W, v_ij, Q accelerations,
//     the acceleration increment) only flip sign with the frame, exactly;
//   * products of two odd quantities (mu, v.gradW, dyads, r r) are unchanged;
//   * the two terms that depend on who owns the pair are kept: the Q
//     accelerations enter the acceleration sum owner first, and the tensile
//     factor uses the owner's W.
// Sum density adds the self term where the reference does: after the
// contributions of lower-index neighbours, before those of higher ones.
// With H = h I (every off-diagonal exactly zero) the SymTensor products reduce
// to h*r componentwise with the same bits, which the isotropic kernels use.
//
// Memory: neighbour indices arrive four at a time (one int4 of the blocked ELL),
// the next neighbour's record is loaded while the current pair is evaluated,
// and the kernel table sits in LDS.
#include <type_traits>
#include "internal.hh"
#include "dgeom.hh"
#include "records.hh"
#include "pack.hh"
#include "ell.hh"
#include "eigen.hh"
#include <cstring>
#include <cmath>

namespace sphgpu {

namespace {

constexpr int kBlock = 256;
inline int nblocks(int64_t n, int bs = kBlock) { return int(std::max<int64_t>(1, (n + bs - 1) / bs)); }

// ---------------------------------------------------------------------------
// Record pack: every slot, or the `count` slots listed in `slots`.
// ---------------------------------------------------------------------------
template <int D, bool ISO, int SEC = 0>
__global__ void k_pack(PackIn in, PackOut o, const int32_t* __restrict__ slots, int64_t count) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= count) return;
  packSlot<D, ISO, SEC>(slots ? int64_t(slots[t]) : t, in, o);
}

// Table: per point [Ak Bk Ck Ag Bg Cg].
struct TabArgs {
  const double* g;     // global copy
  const float* g32;    // FP32 copy, 8 floats per point
  int n;
  double step, extent;
  const double* nperh;
  const double* wsum;
};

__device__ __forceinline__ void lookupWG(const double* __restrict__ tab, const TabArgs& t, double eta, double Hdet,
                                         double& W, double& gW) {
  if (eta < t.extent) {
    const double q = eta / t.step;
    int lb = int(q);
    lb = lb < t.n - 1 ? lb : t.n - 1;
    const int i0 = lb < t.n - 3 ? lb : t.n - 3;
    const double x = q - i0;
    const double* c = tab + 6 * (i0 + 1);
    W = Hdet * (c[0] + c[1] * x + c[2] * x * x);
    gW = Hdet * (c[3] + c[4] * x + c[5] * x * x);
  } else {
    W = 0.0; gW = 0.0;
  }
}
__device__ __forceinline__ double lookupW(const double* __restrict__ tab, const TabArgs& t, double eta, double Hdet) {
  if (!(eta < t.extent)) return 0.0;
  const double q = eta / t.step;
  int lb = int(q);
  lb = lb < t.n - 1 ? lb : t.n - 1;
  const int i0 = lb < t.n - 3 ? lb : t.n - 3;
  const double x = q - i0;
  const double* c = tab + 6 * (i0 + 1);
  return Hdet * (c[0] + c[1] * x + c[2] * x * x);
}

// Largest table that is staged in LDS (the 1000-point default takes 48 KB).
constexpr int kLdsTablePoints = 1365;
// Staging the table costs a full table read and a barrier per workgroup; it
// pays when each node makes many lookups (2-D rows hold ~28 entries, 3-D rows
// ~120), not for 1-D rows of ~6, which read the table through L1.
constexpr double kLdsMinRow = 16.0;
bool ldsTable(const Sim& s) {
  if (s.table.n > kLdsTablePoints) return false;
  const double perNode = s.nInt > 0 ? double(s.nEntries) / double(s.nInt) : 0.0;
  return perNode >= kLdsMinRow;
}

template <bool LDS> __device__ __forceinline__ const double* stageTable(double* lds, const TabArgs& t) {
  if constexpr (LDS) {
    for (int q = threadIdx.x; q < 6 * t.n; q += blockDim.x) lds[q] = t.g[q];
    __syncthreads();
    return lds;
  } else {
    return t.g;
  }
}

// TableKernel::equivalentNodesPerSmoothingScale (TableKernel.cc:399-429).
__device__ double equivalentNperh(const TabArgs& t, double Wsum) {
  const int n = t.n;
  const double* ws = t.wsum;
  const double* np = t.nperh;
  const bool ascnd = ws[n - 1] >= ws[0];
  int jl = -1, ju = n;
  while (ju - jl > 1) {
    const int jm = (ju + jl) / 2;
    if ((Wsum >= ws[jm]) == ascnd) jl = jm; else ju = jm;
  }
  const int lb = jl, ub = jl + 1;
  if (lb == -1) return np[0];
  if (ub == n) return np[n - 1];
  return dmin(np[ub], dmax(np[lb], np[lb] + (Wsum - ws[lb]) / (ws[ub] - ws[lb]) * (np[ub] - np[lb])));
}

// SPHSmoothingScale::idealSmoothingScale (src/NodeList/SPHSmoothingScale.cc:120-197):
// 1/h1 of the ideal H = h1^-1 I from the root-nu zeroth moment and H.xx.
__device__ __forceinline__ double idealHinv(const TabArgs& t, double zeroth, double Hxx, double nPerh, double hmin,
                                            double hmax) {
  double nperhNow;
  if (fuzzyEqual(zeroth, 0.0)) nperhNow = 0.5 * nPerh;
  else nperhNow = equivalentNperh(t, zeroth);
  const double s = dmin(4.0, dmax(0.25, nPerh / (nperhNow + 1.0e-30)));
  const double aa = (s < 1.0) ? 0.4 * (1.0 + s * s) : 0.4 * (1.0 + 1.0 / (s * s * s));
  const double hi0 = 1.0 / Hxx;
  const double hi1 = dmin(hmax, dmax(hmin, hi0 * (1.0 - aa + aa * s)));
  return 1.0 / hi1;
}

// H.v: with isotropic H the product is h*v componentwise (same bits as the
// SymTensor product with zero off-diagonals).
template <int D, bool ISO> __device__ __forceinline__ V<D> hdot(const S<D>& H, const V<D>& v) {
  if constexpr (ISO) {
    V<D> o;
    for (int c = 0; c < D; ++c) o.c[c] = H.c[0] * v.c[c];
    return o;
  } else {
    return sdot(H, v);
  }
}

// Neighbour indices of one row arrive four at a time (one int4 of the blocked
// ELL); the next int4 is fetched one chunk ahead.
__device__ __forceinline__ int32_t pick4(const int4 v, int q) {
  return q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w));
}
// Advance to entry w (w >= 1, called for w = 1, 2, ... in order).
__device__ __forceinline__ int32_t nextIndex(const int4* __restrict__ nb4, int64_t n, int64_t k, int len, int w,
                                             int4& cur, int4& nxt) {
  if ((w & 3) == 0) {
    cur = nxt;
    if (w + 4 < len) nxt = nb4[ellChunk((w >> 2) + 1, n, k)];
  }
  return pick4(cur, w & 3);
}

// What one side of a pair reads from a neighbour's record.
template <int D, bool ISO> struct Nbr {
  V<D> x, v;
  S<D> H;
  double hdet, m, orig, rho, cs, prho, mor, p, som, fcl, fcq, fsh;
};
template <int D, bool ISO, bool WANT_P>
__device__ __forceinline__ Nbr<D, ISO> loadNbr(const double* __restrict__ rec, int32_t j) {
  using R = RL<ISO>;
  const double* r = rec + int64_t(j) * R::STRIDE;
  Nbr<D, ISO> o;
  for (int c = 0; c < D; ++c) { o.x.c[c] = r[R::X + c]; o.v.c[c] = r[R::V + c]; }
  if constexpr (ISO) {
    o.H.c[0] = r[RL<true>::HINV];
    for (int c = 1; c < NSym<D>::value; ++c) o.H.c[c] = 0.0;
    for (int c = 1; c < D; ++c) o.H.c[sdiag<D>(c)] = o.H.c[0];
    o.fcl = o.fcq = o.fsh = 1.0;
  } else {
    for (int c = 0; c < NSym<D>::value; ++c) o.H.c[c] = r[RL<false>::H + c];
    o.fcl = r[RL<false>::FCL]; o.fcq = r[RL<false>::FCQ]; o.fsh = r[RL<false>::FSH];
  }
  o.hdet = r[R::HDET]; o.m = r[R::M]; o.orig = r[R::ORIG]; o.rho = r[R::RHO]; o.cs = r[R::CS];
  o.prho = r[R::PRHO]; o.mor = r[R::MOR];
  if constexpr (WANT_P) { o.p = r[R::P]; o.som = r[R::SOM]; } else { o.p = 0.0; o.som = 0.0; }
  return o;
}

// ---------------------------------------------------------------------------
// computeSPHSumMassDensity
// ---------------------------------------------------------------------------
template <int D, bool ISO> struct DNbr {
  V<D> x;
  S<D> H;
  double hdet, m, orig;
};
template <int D, bool ISO>
__device__ __forceinline__ DNbr<D, ISO> loadDNbr(const double* __restrict__ rec, int32_t j) {
  using R = RL<ISO>;
  const double* r = rec + int64_t(j) * R::STRIDE;
  DNbr<D, ISO> o;
  for (int c = 0; c < D; ++c) o.x.c[c] = r[R::X + c];
  if constexpr (ISO) {
    o.H.c[0] = r[RL<true>::HINV];
    for (int c = 1; c < NSym<D>::value; ++c) o.H.c[c] = 0.0;
    for (int c = 1; c < D; ++c) o.H.c[sdiag<D>(c)] = o.H.c[0];
  } else {
    for (int c = 0; c < NSym<D>::value; ++c) o.H.c[c] = r[RL<false>::H + c];
  }
  o.hdet = r[R::HDET]; o.m = r[R::M]; o.orig = r[R::ORIG];
  return o;
}
// A density record (DensRec).  det H of an exactly isotropic H is (1/h)^D as
// sdet evaluates it (the off-diagonal products are exact zeros): bitwise the
// packed value.
template <int D>
__device__ __forceinline__ DNbr<D, true> loadDensRec(const double* __restrict__ recS, int32_t j) {
  const double2* p = reinterpret_cast<const double2*>(recS) + int64_t(j) * 3;
  const double2 a = p[0], b = p[1], c = p[2];
  DNbr<D, true> o;
  o.x.c[0] = a.x;
  if constexpr (D > 1) o.x.c[1] = a.y;
  if constexpr (D > 2) o.x.c[2] = b.x;
  const double hinv = b.y;
  o.H.c[0] = hinv;
  for (int c2 = 1; c2 < NSym<D>::value; ++c2) o.H.c[c2] = 0.0;
  for (int c2 = 1; c2 < D; ++c2) o.H.c[sdiag<D>(c2)] = hinv;
  o.hdet = D == 1 ? hinv : (D == 2 ? hinv * hinv : (hinv * hinv) * hinv);
  o.m = c.x; o.orig = c.y;
  return o;
}

// Neighbours come from nrec: the density records (DensRec) for isotropic H, else
// the full records.
template <int D, bool ISO, bool LDS>
__global__ void __launch_bounds__(kBlock) k_sum_density(double* __restrict__ rec, const int32_t* __restrict__ rowCount,
                                                        const int32_t* __restrict__ nbr, int64_t n, int64_t nInt,
                                                        const int32_t* __restrict__ work, TabArgs t,
                                                        double* __restrict__ rhoOut, const double* __restrict__ nrec,
                                                        const double* __restrict__ listOf, int sumAll) {
  using R = RL<ISO>;
  extern __shared__ double lds[];
  const double* tab = stageTable<LDS>(lds, t);
  const int64_t tw = xcdBlock((nInt + kBlock - 1) / kBlock) * kBlock + threadIdx.x;
  if (tw >= nInt) return;
  const int64_t k = work[tw];
  double* ri_ = rec + k * R::STRIDE;
  const double oi = ri_[R::ORIG];
  if (oi >= double(nInt)) return;
  const DNbr<D, ISO> me = loadDNbr<D, ISO>(rec, int32_t(k));
  const double mi = me.m;
  const double self = mi * lookupW(tab, t, 0.0, me.hdet);
  double acc = 0.0;
  bool selfDone = false;
  const int len = rowCount[k];
  const int4* nb4 = reinterpret_cast<const int4*>(nbr);
  int4 cur = len > 0 ? nb4[ellChunk(0, n, k)] : make_int4(0, 0, 0, 0);
  int4 nxt = len > 4 ? nb4[ellChunk(1, n, k)] : make_int4(0, 0, 0, 0);
  DNbr<D, ISO> nx;
  auto load = [&](int32_t j) {
    if constexpr (ISO) return loadDensRec<D>(nrec, j);
    else return loadDNbr<D, false>(nrec, j);
  };
  if (len > 0) nx = load(cur.x);
  for (int w = 0; w < len; ++w) {
    const DNbr<D, ISO> nb = nx;
    if (w + 1 < len) nx = load(nextIndex(nb4, n, k, len, w + 1, cur, nxt));
    if (!selfDone && nb.orig > oi) { acc += self; selfDone = true; }
    // several NodeLists: a pair across them counts with sumAll only, the node
    // taking its own mass (computeSPHSumMassDensity.cc:77-100)
    const bool same = !listOf || listOf[int64_t(nb.orig)] == listOf[int64_t(oi)];
    if (!same && !sumAll) continue;
    const V<D> rij = vsub(me.x, nb.x);
    const double etaj = vmag(hdot<D, ISO>(nb.H, rij));
    acc += (same ? nb.m : mi) * lookupW(tab, t, etaj, nb.hdet);
  }
  if (!selfDone) acc += self;
  rhoOut[int64_t(oi)] = acc;
  // keep the record consistent with the new density (the pass reads only
  // positions, H and masses of the neighbours)
  const double som = ri_[R::SOM], p = ri_[R::P];
  ri_[R::RHO] = acc;
  ri_[R::PRHO] = som * p / (acc * acc);
  ri_[R::MOR] = mi / acc;
}

// ---------------------------------------------------------------------------
// computeSPHOmegaGradhCorrection (src/SPH/computeSPHOmegaGradhCorrection.cc:26-133)
// Gather form of the owner loop: node i receives W(|H_i r_ij|) det H_i and
// |H_i r_ij| W' from every neighbour in row order, its self term W(0) det H_i
// after the lower-index neighbours, and omega_i = -sum(eta W')/(nu sum W); a
// node without neighbours gets 1.  With index-ordered rows this is the
// reference's sum order; with several NodeLists the reference-order rows come
// in the order the owner loop delivers node i's terms (the internal owners of
// lower index first, then i's own pairs), and the owner of a pair is still its
// end of lower caller index, so the self term goes to the same place.
// ---------------------------------------------------------------------------
template <int D, bool ISO, bool LDS>
__global__ void __launch_bounds__(kBlock) k_omega(const double* __restrict__ rec, const int32_t* __restrict__ rowCount,
                                                  const int32_t* __restrict__ nbr, int64_t n, int64_t nInt,
                                                  const int32_t* __restrict__ work, TabArgs t,
                                                  double* __restrict__ omegaOut) {
  extern __shared__ double lds[];
  const double* tab = stageTable<LDS>(lds, t);
  const int64_t tw = xcdBlock((nInt + kBlock - 1) / kBlock) * kBlock + threadIdx.x;
  if (tw >= nInt) return;
  const int64_t k = work[tw];
  const DNbr<D, ISO> me = loadDNbr<D, ISO>(rec, int32_t(k));
  const int len = rowCount[k];
  if (len == 0) {
    omegaOut[int64_t(me.orig)] = 1.0;
    return;
  }
  const double self = lookupW(tab, t, 0.0, me.hdet);
  double om = 0.0, gs = 0.0;
  bool selfDone = false;
  const int4* nb4 = reinterpret_cast<const int4*>(nbr);
  int4 cur = nb4[ellChunk(0, n, k)];
  int4 nxt = len > 4 ? nb4[ellChunk(1, n, k)] : make_int4(0, 0, 0, 0);
  DNbr<D, ISO> nx = loadDNbr<D, ISO>(rec, cur.x);
  for (int w = 0; w < len; ++w) {
    const DNbr<D, ISO> nb = nx;
    if (w + 1 < len) nx = loadDNbr<D, ISO>(rec, nextIndex(nb4, n, k, len, w + 1, cur, nxt));
    if (!selfDone && nb.orig > me.orig) { om += self; selfDone = true; }
    const V<D> rij = vsub(me.x, nb.x);
    const double eta = vmag(hdot<D, ISO>(me.H, rij));
    double Wv, gW;
    lookupWG(tab, t, eta, me.hdet, Wv, gW);
    om += Wv;
    gs += eta * gW;
  }
  if (!selfDone) om += self;
  omegaOut[int64_t(me.orig)] = -gs / (double(D) * om);
}

// ---------------------------------------------------------------------------
// evaluateDerivatives, FP64
// ---------------------------------------------------------------------------
struct PassArgs {
  double Cl, Cq, eps2, epsTensile, nPerh, hmin, hmax;
  double W0, WnPerh;
  int XSPH, correctVelocityGradient, evolveTotalEnergy, linearInExpansion, quadraticInExpansion;
  double qInv;   // 1/u and u of the 3-D mixed record's positions
  float qUnit;
  const double* listOf;   // NodeList of every caller row (several NodeLists), or null
  const int32_t* slotList;   // NodeList of every slot (several NodeLists, collapsed kernel), or null
};

__device__ __forceinline__ double muQ(double mu, double Cl, double Cq, double cs, const PassArgs& p) {
  const double m0 = dmin(0.0, mu);
  const double lin = p.linearInExpansion ? mu : m0;
  const double quad = p.quadraticInExpansion ? -(mu < 0.0 ? -1.0 : 1.0) * mu * mu : m0 * m0;
  return -Cl * cs * lin + Cq * quad;
}

struct OutArgs {
  double* base;      // field-major blocks, each n rows
  int64_t n;         // rows per block (sorted slots)
  int off[SPHGPU_NUM_DERIV_FIELDS];
  int updateOnly;    // store only the fields an integrator update reads
  double* pairDv;    // compatible energy: the pair term of every row entry in the node's own
                     // frame, component c at pairDv[c*plane + ellIndex(w, n, k)]; null otherwise
  int64_t plane;
};

__device__ __forceinline__ double* outRow(const OutArgs& o, int field, int64_t k, int comps) {
  return o.base + int64_t(o.off[field]) * o.n + k * comps;
}

// Per-node accumulators of the pair loop (FP64 in both precisions).
template <int D> struct Acc {
  double rhoSum, norm, DepsDt, maxVP, effVP, viscW, xsphW, wNS;
  V<D> DvDt, xsphDV;
  T<D> DvDx, M;
  S<D> m2nd;
  __device__ __forceinline__ void zero() {
    rhoSum = norm = DepsDt = maxVP = effVP = viscW = xsphW = wNS = 0.0;
    DvDt = vzero<D>(); xsphDV = vzero<D>(); DvDx = tzero<D>(); M = tzero<D>(); m2nd = szero<D>();
  }
};

// Own-node values the epilogue reads.
template <int D> struct Self {
  V<D> v;
  S<D> H;
  double Hdet, m, rho, mor;
};

// Per-node epilogue (SPHHydroBase.cc:834-902) and the output stores.
template <int D>
__device__ __forceinline__ void epilogue(int64_t k, int len, const Self<D>& me, Acc<D>& a, const TabArgs& t,
                                         const PassArgs& p, const OutArgs& o) {
  constexpr int NS = NSym<D>::value;
  const double tiny = 1.0e-30;
  const V<D> vi = me.v;
  const S<D> Hi = me.H;
  const double Hdeti = me.Hdet, mi = me.m, rhoi = me.rho, mori = me.mor;
  double rhoSum = a.rhoSum, norm = a.norm, DepsDt = a.DepsDt, xsphW = a.xsphW;
  V<D> DvDt = a.DvDt, xsphDV = a.xsphDV;
  T<D> DvDx = a.DvDx, M = a.M;
  S<D> m2nd = a.m2nd;
  rhoSum += mi * p.W0 * Hdeti;
  norm += mori * p.W0 * Hdeti;
  T<D> Mout = M;
  if (p.correctVelocityGradient && fabs(tdet(M)) > 1.0e-10 && double(len) > pownu2<D>()) {
    Mout = tinv(M);
    DvDx = tdot(DvDx, Mout);
  } else {
    tdivEq(DvDx, rhoi);
  }
  const double trace = ttrace(DvDx);
  const double DrhoDt = -rhoi * trace;
  if (p.evolveTotalEnergy) DepsDt = mi * (vdot(vi, DvDt) + DepsDt);
  const double wNSo = rootnu<D>(dmax(0.0, a.wNS / Hdeti));
  m2nd = sdivs(m2nd, Hdeti * Hdeti);
  V<D> DxDt = vi;
  if (p.XSPH) {
    xsphW += Hdeti * mi / rhoi * p.W0;
    const V<D> corr = vdivs(xsphDV, dmax(tiny, xsphW));
    for (int c = 0; c < D; ++c) DxDt.c[c] = vi.c[c] + corr.c[c];
  }
  // SPHSmoothingScale::smoothingScaleDerivative: -H/nu * tr(DvDx)
  S<D> negH;
  for (int c = 0; c < NS; ++c) negH.c[c] = -Hi.c[c];
  S<D> DHDt = sdivs(negH, double(D));
  for (int c = 0; c < NS; ++c) DHDt.c[c] = DHDt.c[c] * trace;
  // SPHSmoothingScale::idealSmoothingScale
  const double hinv = idealHinv(t, wNSo, Hi.c[0], p.nPerh, p.hmin, p.hmax);

  // outputs
  // the fields an integrator update reads
  *outRow(o, SPHGPU_D_DRHODT, k, 1) = DrhoDt;
  *outRow(o, SPHGPU_D_DEPSDT, k, 1) = DepsDt;
  {
    double* dx = outRow(o, SPHGPU_D_DXDT, k, D);
    double* dv = outRow(o, SPHGPU_D_DVDT, k, D);
    for (int c = 0; c < D; ++c) { dx[c] = DxDt.c[c]; dv[c] = DvDt.c[c]; }
    double* hid = outRow(o, SPHGPU_D_HIDEAL, k, NS);
    for (int a2 = 0, s2 = 0; a2 < D; ++a2)
      for (int b2 = a2; b2 < D; ++b2, ++s2) hid[s2] = (a2 == b2 ? 1.0 : 0.0) * hinv;
  }
  if (o.updateOnly) return;
  *outRow(o, SPHGPU_D_RHO_SUM, k, 1) = rhoSum;
  *outRow(o, SPHGPU_D_NORMALIZATION, k, 1) = norm;
  *outRow(o, SPHGPU_D_MAX_VISCOUS_PRESSURE, k, 1) = a.maxVP;
  *outRow(o, SPHGPU_D_EFF_VISCOUS_PRESSURE, k, 1) = a.effVP;
  *outRow(o, SPHGPU_D_VISCOUS_WORK, k, 1) = a.viscW;
  *outRow(o, SPHGPU_D_XSPH_WEIGHT_SUM, k, 1) = xsphW;
  *outRow(o, SPHGPU_D_WEIGHTED_NEIGHBOR_SUM, k, 1) = wNSo;
  {
    double* xv = outRow(o, SPHGPU_D_XSPH_DELTA_V, k, D);
    for (int c = 0; c < D; ++c) xv[c] = xsphDV.c[c];
  }
  {
    double* g = outRow(o, SPHGPU_D_DVDX, k, D * D);
    double* gl = outRow(o, SPHGPU_D_LOCAL_DVDX, k, D * D);
    double* m = outRow(o, SPHGPU_D_M, k, D * D);
    double* ml = outRow(o, SPHGPU_D_LOCAL_M, k, D * D);
    for (int c = 0; c < D * D; ++c) { g[c] = DvDx.c[c]; gl[c] = DvDx.c[c]; m[c] = Mout.c[c]; ml[c] = Mout.c[c]; }
  }
  {
    double* dh = outRow(o, SPHGPU_D_DHDT, k, NS);
    double* m2 = outRow(o, SPHGPU_D_MASS_SECOND_MOMENT, k, NS);
    for (int c = 0; c < NS; ++c) { dh[c] = DHDt.c[c]; m2[c] = m2nd.c[c]; }
  }
}

template <int D, bool ISO, bool TENSILE>
__device__ __forceinline__ unsigned derivsNodeFP64(int64_t k, const double* __restrict__ rec,
                                                   const int32_t* __restrict__ rowCount, const int32_t* __restrict__ nbr,
                                                   int64_t n, const double* __restrict__ tab, const TabArgs& t,
                                                   const PassArgs& p, const OutArgs& o) {
  const Nbr<D, ISO> me = loadNbr<D, ISO, true>(rec, int32_t(k));
  const V<D> ri = me.x, vi = me.v;
  const S<D> Hi = me.H;
  const double oi = me.orig;
  const double Hdeti = me.hdet, mi = me.m, rhoi = me.rho, ci = me.cs, Pi = me.p;
  const double somi = me.som, Prhoi0 = me.prho, mori = me.mor;

  double rhoSum = 0.0, norm = 0.0, DepsDt = 0.0, maxVP = 0.0, effVP = 0.0, viscW = 0.0, xsphW = 0.0, wNS = 0.0;
  V<D> DvDt = vzero<D>(), xsphDV = vzero<D>();
  T<D> DvDx = tzero<D>(), M = tzero<D>();
  S<D> m2nd = szero<D>();
  unsigned pairs = 0;

  const int len = rowCount[k];
  const int4* nb4 = reinterpret_cast<const int4*>(nbr);
  int4 cur = len > 0 ? nb4[ellChunk(0, n, k)] : make_int4(0, 0, 0, 0);
  int4 nxt = len > 4 ? nb4[ellChunk(1, n, k)] : make_int4(0, 0, 0, 0);
  Nbr<D, ISO> nx;
  if (len > 0) nx = loadNbr<D, ISO, TENSILE>(rec, cur.x);
  for (int w = 0; w < len; ++w) {
    const Nbr<D, ISO> nb = nx;
    if (w + 1 < len) nx = loadNbr<D, ISO, TENSILE>(rec, nextIndex(nb4, n, k, len, w + 1, cur, nxt));
    const bool ownerI = nb.orig > oi;
    pairs += ownerI ? 1u : 0u;
    const V<D> rij = vsub(ri, nb.x);
    const V<D> vij = vsub(vi, nb.v);
    const S<D>& Hj = nb.H;
    const double Hdetj = nb.hdet, mj = nb.m, rhoj = nb.rho, cj = nb.cs, morj = nb.mor;

    const V<D> etai = hdot<D, ISO>(Hi, rij), etaj = hdot<D, ISO>(Hj, rij);
    const double eta2i = vmag2(etai), eta2j = vmag2(etaj);
    const double etaMagi = (D == 1) ? fabs(etai.c[0]) : __builtin_sqrt(eta2i);
    const double etaMagj = (D == 1) ? fabs(etaj.c[0]) : __builtin_sqrt(eta2j);
    const V<D> Hetai = hdot<D, ISO>(Hi, vunit(etai, etaMagi));
    const V<D> Hetaj = hdot<D, ISO>(Hj, vunit(etaj, etaMagj));
    double Wi, gWi, Wj, gWj;
    lookupWG(tab, t, etaMagi, Hdeti, Wi, gWi);
    lookupWG(tab, t, etaMagj, Hdetj, Wj, gWj);
    const V<D> gradWi = vscale(Hetai, gWi);
    const V<D> gradWj = vscale(Hetaj, gWj);

    // moments for the ideal H (SPHHydroBase.cc:737-743)
    const double rij2 = vmag2(rij);
    const double sinv = safeInvVar(rij2 * rij2 * rij2);
    const double g2 = vmag2(gradWi);
    wNS += fabs(gWi);
    {
      int s = 0;
      for (int a = 0; a < D; ++a)
        for (int b = a; b < D; ++b, ++s) m2nd.c[s] += (rij.c[a] * rij.c[b]) * sinv * g2;
    }
    if (!p.listOf || p.listOf[int64_t(nb.orig)] == p.listOf[int64_t(oi)]) {   // SPHHydroBase.cc:746-751
      rhoSum += mj * Wi;
      norm += mori * Wi;
    }

    // Monaghan-Gingold Q (MonaghanGingoldViscosity.cc:83-132)
    double Cl = p.Cl, Cq = p.Cq;
    if constexpr (!ISO) {
      const double fsh = dmax(me.fsh, nb.fsh);
      Cl *= 0.5 * (me.fcl + nb.fcl) * fsh;
      Cq *= 0.5 * (me.fcq + nb.fcq) * fsh;
    }
    const double mui = vdot(vij, etai) / (eta2i + p.eps2);
    const double muj = vdot(vij, etaj) / (eta2j + p.eps2);
    const double QPi = muQ(mui, Cl, Cq, ci, p) / rhoi;
    const double QPj = muQ(muj, Cl, Cq, cj, p) / rhoj;
    V<D> Qacci, Qaccj;
    for (int c = 0; c < D; ++c) { Qacci.c[c] = 0.5 * (QPi * gradWi.c[c]); Qaccj.c[c] = 0.5 * (QPj * gradWj.c[c]); }
    const double workQi = vdot(vij, Qacci);
    const double Qi = rhoi * rhoi * fabs(QPi);
    maxVP = dmax(maxVP, Qi);
    effVP += morj * Qi * Wi;
    viscW += mj * workQi;

    double Prhoi = Prhoi0, Prhoj = nb.prho;
    if constexpr (TENSILE) {
      const double Wo = ownerI ? Wi : Wj, Hdeto = ownerI ? Hdeti : Hdetj;
      const double x = Wo / (Hdeto * p.WnPerh);
      const double fij = p.epsTensile * (x * x * x * x);
      const double Pj = nb.p;
      Prhoi = somi * (Pi + fij * (Pi < 0.0 ? -Pi : 0.0)) / (rhoi * rhoi);
      Prhoj = nb.som * (Pj + fij * (Pj < 0.0 ? -Pj : 0.0)) / (rhoj * rhoj);
    }
    for (int c = 0; c < D; ++c) {
      const double qf = ownerI ? Qacci.c[c] : Qaccj.c[c];
      const double qs = ownerI ? Qaccj.c[c] : Qacci.c[c];
      const double dv = gradWi.c[c] * Prhoi + gradWj.c[c] * Prhoj + qf + qs;
      DvDt.c[c] += dv * (-mj);
      if (o.pairDv) o.pairDv[int64_t(c) * o.plane + ellIndex(w, n, k)] = dv;
    }
    DepsDt += mj * (Prhoi * vdot(vij, gradWi) + workQi);
    for (int a = 0; a < D; ++a)
      for (int b = 0; b < D; ++b) {
        DvDx.c[a * D + b] -= (vij.c[a] * gradWi.c[b]) * mj;
        M.c[a * D + b] -= (rij.c[a] * gradWi.c[b]) * mj;
      }
    if (p.XSPH) {
      const double w2 = 0.5 * (mori * Wi + morj * Wj);
      xsphW += w2;
      for (int c = 0; c < D; ++c) xsphDV.c[c] -= vij.c[c] * w2;
    }
  }

  // per-node epilogue (SPHHydroBase.cc:834-902)
  Acc<D> a;
  a.rhoSum = rhoSum; a.norm = norm; a.DepsDt = DepsDt; a.maxVP = maxVP; a.effVP = effVP; a.viscW = viscW;
  a.xsphW = xsphW; a.wNS = wNS; a.DvDt = DvDt; a.xsphDV = xsphDV; a.DvDx = DvDx; a.M = M; a.m2nd = m2nd;
  Self<D> self;
  self.v = vi; self.H = Hi; self.Hdet = Hdeti; self.m = mi; self.rho = rhoi; self.mor = mori;
  epilogue<D>(k, len, self, a, t, p, o);
  return pairs;
}

template <int D, bool ISO, bool TENSILE, bool LDS>
__global__ void __launch_bounds__(kBlock) k_derivs_fp64(const double* __restrict__ rec, const int32_t* __restrict__ rowCount,
                                                        const int32_t* __restrict__ nbr, int64_t n, int64_t nInt,
                                                        const int32_t* __restrict__ work, TabArgs t, PassArgs p,
                                                        OutArgs o, unsigned long long* __restrict__ ncalc) {
  extern __shared__ double lds[];
  const double* tab = stageTable<LDS>(lds, t);
  const int64_t tw = xcdBlock((nInt + kBlock - 1) / kBlock) * kBlock + threadIdx.x;
  unsigned pairs = 0;
  if (tw < nInt) pairs = derivsNodeFP64<D, ISO, TENSILE>(work[tw], rec, rowCount, nbr, n, tab, t, p, o);
  for (int sh = 32; sh > 0; sh >>= 1) pairs += __shfl_xor(pairs, sh);
  if ((threadIdx.x & 63) == 0 && pairs) atomicAdd(ncalc, (unsigned long long)pairs);
}

// ---------------------------------------------------------------------------
// Isotropic pair loop in collapsed form, FP32 (mixed precision) or FP64 (the
// search-order FP64 path).  Positions and per-node accumulators are FP64.
//
// With H = h I both unit vectors of a pair are r_hat = r_ij/|r_ij| and
// grad W = r_hat * h * dW/deta * det H, so the pair body reduces to scalars
// along r_hat: the acceleration increment is r_hat * m_j * (g_i (P_i/rho_i^2 +
// Q_i/2) + g_j (P_j/rho_j^2 + Q_j/2)) with g = h W' det H.  The two ends of a
// pair compute the same scalar from the same inputs (r_ij is the negated FP64
// difference, r_hat flips sign exactly), so the pair forces are antisymmetric
// up to the final product with the partner mass.  This reorders the
// reference's arithmetic: in FP64 the results agree with the reference-order
// kernel to round-off (the bar for search row order is 1e-12), in FP32 to the
// mixed-precision bar (1e-5).
//
// Neighbour records (written in pack.hh): the mixed record (MixedRec, 32 / 48 /
// 48 B in 1-D / 2-D / 3-D) and the FP64 record (96 B in two planes); both mark
// ghosts by a negative 1/h, and the reader forms det H = (1/h)^D.
// ---------------------------------------------------------------------------
template <typename Real> struct INbr {
  double x[3];
  uint32_t q[3];    // 3-D mixed: quantised position words
  Real hinv, hdet, m, invRho, v[3], cs, prho, mor;
  bool ghost;
};
template <int D>
__device__ __forceinline__ void loadIso(const float4* __restrict__ recM, int64_t, int32_t j, INbr<float>& o) {
  const float4* p = recM + int64_t(j) * MixedRec<D>::F4;
  auto dbl = [](float lo, float hi) { return __hiloint2double(__float_as_int(hi), __float_as_int(lo)); };
  float h;
  if constexpr (D == 1) {
    const float4 f0 = p[0], f1 = p[1];
    o.x[0] = dbl(f0.x, f0.y); o.x[1] = 0.0; o.x[2] = 0.0;
    h = f0.z; o.m = f0.w;
    o.invRho = f1.x; o.v[0] = f1.y; o.v[1] = 0.0f; o.v[2] = 0.0f; o.cs = f1.z; o.prho = f1.w;
    o.mor = o.m * o.invRho;
  } else if constexpr (D == 2) {
    const float4 f0 = p[0], f1 = p[1], f2 = p[2];
    o.x[0] = dbl(f0.x, f0.y); o.x[1] = dbl(f0.z, f0.w); o.x[2] = 0.0;
    h = f1.x; o.m = f1.y; o.invRho = f1.z; o.v[0] = f1.w;
    o.v[1] = f2.x; o.v[2] = 0.0f; o.cs = f2.y; o.prho = f2.z; o.mor = f2.w;
  } else {
    const float4 f0 = p[0], f1 = p[1], f2 = p[2];
    o.q[0] = __float_as_uint(f0.x); o.q[1] = __float_as_uint(f0.y); o.q[2] = __float_as_uint(f0.z);
    h = f0.w;
    o.m = f1.x; o.invRho = f1.y; o.v[0] = f1.z; o.v[1] = f1.w;
    o.v[2] = f2.x; o.cs = f2.y; o.prho = f2.z; o.mor = f2.w;
  }
  o.ghost = h < 0.0f;
  o.hinv = fabsf(h);
  o.hdet = D == 1 ? o.hinv : (D == 2 ? o.hinv * o.hinv : o.hinv * o.hinv * o.hinv);
}
// det H of an exactly isotropic H is (1/h)^D evaluated as sdet evaluates it
// (the off-diagonal products are exact zeros), so the value is bitwise the
// packed one; m/rho is m * (1/rho), within an ulp.
template <int D>
__device__ __forceinline__ void loadIso(const double2* __restrict__ recD, int64_t n, int32_t j, INbr<double>& o) {
  const double2* a = recD + int64_t(j) * 4;
  const double2* b = recD + 4 * n + int64_t(j) * 2;
  const double2 d0 = a[0], d1 = a[1], d2 = a[2], d3 = a[3], e0 = b[0], e1 = b[1];
  o.x[0] = d0.x; o.x[1] = d0.y; o.x[2] = d1.x;
  o.ghost = d1.y < 0.0;
  const double hinv = fabs(d1.y);
  o.hinv = hinv;
  o.hdet = D == 1 ? hinv : (D == 2 ? hinv * hinv : (hinv * hinv) * hinv);
  o.m = d2.x; o.invRho = d2.y; o.cs = d3.x; o.prho = d3.y;
  o.mor = d2.x * d2.y;
  o.v[0] = e0.x; o.v[1] = e0.y; o.v[2] = e1.x;
}

template <int D>
__global__ void k_pack_mixed(const double* __restrict__ rec, int64_t n, int64_t nInt, double qInv,
                             float4* __restrict__ recM) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k < n) writeMixedRec<D>(rec + k * RL<true>::STRIDE, nInt, qInv, recM + k * MixedRec<D>::F4);
}

template <int D>
__global__ void k_pack_fast64(const double* __restrict__ rec, int64_t n, int64_t nInt, double2* __restrict__ recD) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k < n) writeFast64Rec<D>(rec + k * RL<true>::STRIDE, nInt, recD + k * 4, recD + 4 * n + k * 2);
}

template <int D>
__global__ void k_pack_dens_mixed(const double* __restrict__ rec, int64_t n, float4* __restrict__ recSM) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k < n) writeDensRecMixed<D>(rec + k * RL<true>::STRIDE, recSM + k * DensMixedRec<D>::F4);
}

template <int D>
__global__ void k_pack_dens(const double* __restrict__ rec, int64_t n, double2* __restrict__ recS) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k < n) writeDensRec<D>(rec + k * RL<true>::STRIDE, recS + k * 3);
}

// FP32 table in LDS: 8 floats per point.
template <bool LDS> __device__ __forceinline__ const float* stageTable32(float* lds, const TabArgs& t) {
  if constexpr (LDS) {
    for (int q = threadIdx.x; q < 8 * t.n; q += blockDim.x) lds[q] = t.g32[q];
    __syncthreads();
    return lds;
  } else {
    return t.g32;
  }
}
struct Tab32 {
  float invStep, extent;
  int n;
};
struct Tab64 {
  double invStep, extent;
  int n;
};

__device__ __forceinline__ float fmaR(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaR(double a, double b, double c) { return __builtin_fma(a, b, c); }
// reciprocal and reciprocal square root: the hardware approximations in FP32,
// refined by two Newton steps (error squared each step) in FP64
__device__ __forceinline__ float rcpR(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcpR(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = __builtin_fma(y, __builtin_fma(-x, y, 1.0), y);
  return __builtin_fma(y, __builtin_fma(-x, y, 1.0), y);
}
__device__ __forceinline__ float rsqR(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqR(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = __builtin_fma(0.5 * y, __builtin_fma(-(x * y), y, 1.0), y);
  return __builtin_fma(0.5 * y, __builtin_fma(-(x * y), y, 1.0), y);
}
__device__ __forceinline__ float absR(float x) { return fabsf(x); }
__device__ __forceinline__ double absR(double x) { return fabs(x); }
__device__ __forceinline__ float maxR(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double maxR(double a, double b) { return fmax(a, b); }

__device__ __forceinline__ void lookupWGR(const float* __restrict__ tab, const Tab32& t, float eta, float hdet,
                                          float& W, float& gW) {
  if (eta < t.extent) {
    const float q = eta * t.invStep;
    int lb = int(q);
    lb = lb < t.n - 1 ? lb : t.n - 1;
    const int i0 = lb < t.n - 3 ? lb : t.n - 3;
    const float x = q - float(i0);
    const float4 a = *reinterpret_cast<const float4*>(tab + 8 * (i0 + 1));
    const float4 g = *reinterpret_cast<const float4*>(tab + 8 * (i0 + 1) + 4);
    W = hdet * __builtin_fmaf(__builtin_fmaf(a.z, x, a.y), x, a.x);
    gW = hdet * __builtin_fmaf(__builtin_fmaf(g.z, x, g.y), x, g.x);
  } else {
    W = 0.0f; gW = 0.0f;
  }
}
__device__ __forceinline__ void lookupWGR(const double* __restrict__ tab, const Tab64& t, double eta, double hdet,
                                          double& W, double& gW) {
  if (eta < t.extent) {
    const double q = eta * t.invStep;
    int lb = int(q);
    lb = lb < t.n - 1 ? lb : t.n - 1;
    const int i0 = lb < t.n - 3 ? lb : t.n - 3;
    const double x = q - double(i0);
    const double* c = tab + 6 * (i0 + 1);
    W = hdet * __builtin_fma(__builtin_fma(c[2], x, c[1]), x, c[0]);
    gW = hdet * __builtin_fma(__builtin_fma(c[5], x, c[4]), x, c[3]);
  } else {
    W = 0.0; gW = 0.0;
  }
}
__device__ __forceinline__ float lookupW32(const float* __restrict__ tab, const Tab32& t, float eta, float hdet) {
  if (!(eta < t.extent)) return 0.0f;
  const float q = eta * t.invStep;
  int lb = int(q);
  lb = lb < t.n - 1 ? lb : t.n - 1;
  const int i0 = lb < t.n - 3 ? lb : t.n - 3;
  const float x = q - float(i0);
  const float4 a = *reinterpret_cast<const float4*>(tab + 8 * (i0 + 1));
  return hdet * __builtin_fmaf(__builtin_fmaf(a.z, x, a.y), x, a.x);
}
template <typename Real>
__device__ __forceinline__ Real muQR(Real mu, Real Cl, Real Cq, Real cs, int linear, int quadratic) {
  const Real m0 = mu < Real(0) ? mu : Real(0);
  const Real lin = linear ? mu : m0;
  const Real quad = quadratic ? -(mu < Real(0) ? Real(-1) : Real(1)) * mu * mu : m0 * m0;
  return -Cl * cs * lin + Cq * quad;
}

template <int D, bool LDS>
__global__ void __launch_bounds__(kBlock) k_sum_density_mixed(double* __restrict__ rec, const float4* __restrict__ recSM,
                                                              const int32_t* __restrict__ rowCount,
                                                              const int32_t* __restrict__ nbr, int64_t n, int64_t nInt,
                                                              const int32_t* __restrict__ work, TabArgs t, Tab32 t32,
                                                              double W0, double* __restrict__ rhoOut,
                                                              const int32_t* __restrict__ slotList, int sumAll) {
  using R = RL<true>;
  extern __shared__ float lds32[];
  const float* tab = stageTable32<LDS>(lds32, t);
  const int64_t tw = xcdBlock((nInt + kBlock - 1) / kBlock) * kBlock + threadIdx.x;
  if (tw >= nInt) return;
  const int64_t k = work[tw];
  double* ri_ = rec + k * R::STRIDE;
  double xi[3] = {ri_[R::X], D > 1 ? ri_[R::X + 1] : 0.0, D > 2 ? ri_[R::X + 2] : 0.0};
  const double mi = ri_[R::M], hdeti = ri_[R::HDET];
  double acc = mi * (W0 * hdeti);
  const int len = rowCount[k];
  const int4* nb4 = reinterpret_cast<const int4*>(nbr);
  int4 cur = len > 0 ? nb4[ellChunk(0, n, k)] : make_int4(0, 0, 0, 0);
  int4 nxt = len > 4 ? nb4[ellChunk(1, n, k)] : make_int4(0, 0, 0, 0);
  // the mixed density record: x (FP64 bits), 1/h, m
  auto load = [&](int32_t j, float4& a, float4& b) {
    const float4* p = recSM + int64_t(j) * DensMixedRec<D>::F4;
    a = p[0];
    if constexpr (D > 1) b = p[1];
    else b = make_float4(0.0f, 0.0f, a.z, a.w);
  };
  // several NodeLists: a pair across NodeLists counts with sumAll only, the
  // node taking its own mass (computeSPHSumMassDensity.cc:77-100)
  const int32_t lk = slotList ? slotList[k] : 0;
  const float mif = float(mi);
  float4 na, nb2;
  int32_t jn = cur.x;
  if (len > 0) load(jn, na, nb2);
  for (int w = 0; w < len; ++w) {
    const float4 a = na, b = nb2;
    const int32_t j = jn;
    if (w + 1 < len) {
      jn = nextIndex(nb4, n, k, len, w + 1, cur, nxt);
      load(jn, na, nb2);
    }
    float mj = b.w;
    if (slotList && slotList[j] != lk) {
      if (!sumAll) continue;
      mj = mif;
    }
    const double xj[3] = {__hiloint2double(__float_as_int(a.y), __float_as_int(a.x)),
                          __hiloint2double(__float_as_int(a.w), __float_as_int(a.z)),
                          __hiloint2double(__float_as_int(b.y), __float_as_int(b.x))};
    float r2 = 0.0f;
    for (int c = 0; c < D; ++c) {
      const float rc = float(xi[c] - xj[c]);
      r2 = __builtin_fmaf(rc, rc, r2);
    }
    const float hj = b.z;
    const float hdetj = D == 1 ? hj : (D == 2 ? hj * hj : hj * hj * hj);
    const float eta = hj * __builtin_sqrtf(r2);
    acc += double(mj * lookupW32(tab, t32, eta, hdetj));
  }
  rhoOut[int64_t(ri_[R::ORIG])] = acc;
  const double som = ri_[R::SOM], p = ri_[R::P];
  ri_[R::RHO] = acc;
  ri_[R::PRHO] = som * p / (acc * acc);
  ri_[R::MOR] = mi / acc;
}

// The table of the collapsed pair loop: FP32 (8 floats per point) or FP64 (6
// doubles per point), staged in LDS.
template <typename Real> struct IsoTab;
template <> struct IsoTab<float> {
  using T = Tab32;
  using Rec = float4;
  using Elem = float;
  static constexpr int PER = 8;
};
template <> struct IsoTab<double> {
  using T = Tab64;
  using Rec = double2;
  using Elem = double;
  static constexpr int PER = 6;
};

// Mixed precision: the pair terms of up to kSumBlock consecutive neighbours
// are summed in FP32 (fused into the term's last product where it has one),
// and each block's partial sums are added once to the FP64 per-node
// accumulators: one conversion and one FP64 add per block instead of per
// entry.  A block's FP32 sum of 16 terms is within 16 u32 (1e-6) of its
// absolute sum; the per-node sums stay FP64.
constexpr int kSumBlock = 16;
template <int D> struct SumsF {
  float DvDt[D], xsphDV[D], DvDx[D * D], M[D * (D + 1) / 2], m2nd[D * (D + 1) / 2];
  float DepsDt, viscW, effVP, rhoSum, norm, wNS, xsphW;
  __device__ __forceinline__ void zero() {
    for (int c = 0; c < D; ++c) DvDt[c] = xsphDV[c] = 0.0f;
    for (int c = 0; c < D * D; ++c) DvDx[c] = 0.0f;
    for (int c = 0; c < D * (D + 1) / 2; ++c) M[c] = m2nd[c] = 0.0f;
    DepsDt = viscW = effVP = rhoSum = norm = wNS = xsphW = 0.0f;
  }
  // add into the FP64 accumulators (M and m2nd as upper triangles) and restart
  __device__ __forceinline__ void flushInto(Acc<D>& a) {
    for (int c = 0; c < D; ++c) {
      a.DvDt.c[c] += double(DvDt[c]);
      a.xsphDV.c[c] += double(xsphDV[c]);
    }
    for (int c = 0; c < D * D; ++c) a.DvDx.c[c] += double(DvDx[c]);
    int s2 = 0;
    for (int c = 0; c < D; ++c)
      for (int b = c; b < D; ++b, ++s2) {
        a.M.c[c * D + b] += double(M[s2]);
        a.m2nd.c[s2] += double(m2nd[s2]);
      }
    a.DepsDt += double(DepsDt);
    a.viscW += double(viscW);
    a.effVP += double(effVP);
    a.rhoSum += double(rhoSum);
    a.norm += double(norm);
    a.wNS += double(wNS);
    a.xsphW += double(xsphW);
    zero();
  }
};

template <typename Real, int D, bool LDS, bool PREFETCH, int BS = kBlock>
__global__ void __launch_bounds__(BS) k_derivs_iso(const double* __restrict__ rec,
                                                       const typename IsoTab<Real>::Rec* __restrict__ recN,
                                                       const int32_t* __restrict__ rowCount,
                                                       const int32_t* __restrict__ nbr, int64_t n, int64_t nInt,
                                                       const int32_t* __restrict__ work, TabArgs t,
                                                       typename IsoTab<Real>::T tt, PassArgs p, OutArgs o,
                                                       unsigned long long* __restrict__ ncalc) {
  using R = RL<true>;
  using E = typename IsoTab<Real>::Elem;
  extern __shared__ double ldsIso[];
  const E* tab;
  if constexpr (LDS) {
    E* l = reinterpret_cast<E*>(ldsIso);
    const E* g = std::is_same<Real, float>::value ? reinterpret_cast<const E*>(t.g32) : reinterpret_cast<const E*>(t.g);
    for (int q = threadIdx.x; q < IsoTab<Real>::PER * t.n; q += blockDim.x) l[q] = g[q];
    __syncthreads();
    tab = l;
  } else {
    tab = std::is_same<Real, float>::value ? reinterpret_cast<const E*>(t.g32) : reinterpret_cast<const E*>(t.g);
  }
  const int64_t tw = xcdBlock((nInt + BS - 1) / BS) * BS + threadIdx.x;
  unsigned pairs = 0;
  // the 3-D mixed record holds quantised positions
  constexpr bool QPOS = std::is_same<Real, float>::value && D == 3;
  if (tw < nInt) {
    const int64_t k = work[tw];
    const int32_t lk = p.slotList ? p.slotList[k] : 0;
    const double* ri_ = rec + k * R::STRIDE;
    double xi[3];
    uint32_t qi[3];
    Real vi[3];
    for (int c = 0; c < 3; ++c) {
      xi[c] = c < D ? ri_[R::X + c] : 0.0;
      if constexpr (QPOS) qi[c] = quantPos(xi[c], p.qInv);
      vi[c] = c < D ? Real(ri_[R::V + c]) : Real(0);
    }
    const double rho0 = ri_[R::RHO];
    const Real hi = Real(ri_[RL<true>::HINV]), ci = Real(ri_[R::CS]);
    // det H as a neighbour's record gives it (from the same 1/h), so that the
    // two ends of a pair evaluate the same kernel gradient and their pair
    // terms are exact negatives of each other (FP64: the packed det H is
    // already that value)
    const Real hdeti = std::is_same<Real, float>::value ? (D == 1 ? hi : (D == 2 ? hi * hi : hi * hi * hi))
                                                        : Real(ri_[R::HDET]);
    const Real invRhoi = Real(1.0 / rho0), rhoi2 = Real(rho0 * rho0);
    const Real Prhoi = Real(ri_[R::PRHO]), mori = Real(ri_[R::MOR]);
    const Real Cl = Real(p.Cl), Cq = Real(p.Cq), eps2 = Real(p.eps2);

    Acc<D> a;
    a.zero();
    constexpr bool F32 = std::is_same<Real, float>::value;
    SumsF<F32 ? D : 1> af;
    if constexpr (F32) af.zero();
    Real maxVP = Real(0);
    const int len = rowCount[k];
    const int4* nb4 = reinterpret_cast<const int4*>(nbr);
    int4 cur = len > 0 ? nb4[ellChunk(0, n, k)] : make_int4(0, 0, 0, 0);
    int4 nxt = len > 4 ? nb4[ellChunk(1, n, k)] : make_int4(0, 0, 0, 0);
    // One neighbour's contribution, in row order (w: its place in the row).
    auto pair = [&](const INbr<Real>& nb, int32_t j, int w) {
      pairs += (nb.ghost || j > int32_t(k)) ? 1u : 0u;
      Real rij[3], vij[3];
      Real r2 = Real(0), vdotr = Real(0);
      for (int c = 0; c < D; ++c) {
        if constexpr (QPOS) rij[c] = Real(int32_t(qi[c] - nb.q[c])) * p.qUnit;
        else rij[c] = Real(xi[c] - nb.x[c]);
        vij[c] = vi[c] - nb.v[c];
        r2 = fmaR(rij[c], rij[c], r2);
        vdotr = fmaR(vij[c], rij[c], vdotr);
      }
      const Real rinv = r2 > Real(1.0e-30) ? rsqR(r2) : Real(0);
      const Real rmag = r2 * rinv;
      const Real hj = nb.hinv;
      Real Wi, gWi, Wj, gWj;
      lookupWGR(tab, tt, hi * rmag, hdeti, Wi, gWi);
      lookupWGR(tab, tt, hj * rmag, nb.hdet, Wj, gWj);
      const Real gi = hi * gWi, gj = hj * gWj;     // grad W = r_hat * g
      // Monaghan-Gingold Q: mu = v_ij . eta / (eta^2 + eps2), eta = h r_ij
      // (FP64: one refined reciprocal of the product serves both ends)
      const Real da = fmaR(hi * hi, r2, eps2), db = fmaR(hj * hj, r2, eps2);
      Real ia, ib;
      if constexpr (std::is_same<Real, double>::value) {
        const Real inv = rcpR(da * db);
        ia = db * inv;
        ib = da * inv;
      } else {
        ia = rcpR(da);
        ib = rcpR(db);
      }
      const Real mui = hi * vdotr * ia;
      const Real muj = hj * vdotr * ib;
      const Real QPi = muQR(mui, Cl, Cq, ci, p.linearInExpansion, p.quadraticInExpansion) * invRhoi;
      const Real QPj = muQR(muj, Cl, Cq, nb.cs, p.linearInExpansion, p.quadraticInExpansion) * nb.invRho;
      const Real Qi = rhoi2 * absR(QPi);
      maxVP = maxR(maxVP, Qi);
      const Real vdr = vdotr * rinv;                 // v_ij . r_hat
      const Real halfQi = Real(0.5) * QPi * gi;
      const Real pi = fmaR(gi, Prhoi, halfQi);
      const Real pj = fmaR(gj, nb.prho, Real(0.5) * QPj * gj);
      const Real mj = nb.m;
      const Real fsr = mj * (pi + pj) * rinv;        // acceleration along r_ij
      const Real wg = mj * gi * rinv;               // grad W_i * m_j = r_ij * wg
      // gi^2 / max(1e-30, r^6); FP64 takes 1/r^6 from 1/r
      Real f2;
      if constexpr (std::is_same<Real, double>::value) {
        const Real ri2 = rinv * rinv;
        f2 = r2 * r2 * r2 > Real(1.0e-30) ? gi * gi * (ri2 * ri2 * ri2) : gi * gi * Real(1.0e30);
      } else {
        f2 = gi * gi * rcpR(maxR(Real(1.0e-30), r2 * r2 * r2));
      }
      if (o.pairDv) {
        // compatible energy: the own-frame pair term, DvDt_i += -m_j dv
        const Real dsc = (pi + pj) * rinv;
        for (int c = 0; c < D; ++c) o.pairDv[int64_t(c) * o.plane + ellIndex(w, n, k)] = double(rij[c] * dsc);
      }
      // several NodeLists: rhoSum and the normalization take same-NodeList
      // pairs only (SPHHydroBase.cc:746-751)
      const bool sameList = !p.slotList || p.slotList[j] == lk;
      if constexpr (F32) {
        for (int c = 0; c < D; ++c) af.DvDt[c] = fmaR(-rij[c], fsr, af.DvDt[c]);
        af.DepsDt = fmaR(mj * vdr, pi, af.DepsDt);
        af.viscW = fmaR(mj * vdr, halfQi, af.viscW);
        af.effVP = fmaR(nb.mor * Qi, Wi, af.effVP);
        if (sameList) {
          af.rhoSum = fmaR(mj, Wi, af.rhoSum);
          af.norm = fmaR(mori, Wi, af.norm);
        }
        af.wNS += absR(gWi);
        for (int c = 0; c < D; ++c) {
          const Real u = vij[c] * wg;
          for (int b = 0; b < D; ++b) af.DvDx[c * D + b] = fmaR(-u, rij[b], af.DvDx[c * D + b]);
        }
        int s2 = 0;
        for (int c = 0; c < D; ++c) {
          const Real rw = rij[c] * wg, rf = rij[c] * f2;
          for (int b = c; b < D; ++b, ++s2) {
            af.M[s2] = fmaR(-rw, rij[b], af.M[s2]);
            af.m2nd[s2] = fmaR(rf, rij[b], af.m2nd[s2]);
          }
        }
        if (p.XSPH) {
          const Real w2 = Real(0.5) * fmaR(mori, Wi, nb.mor * Wj);
          af.xsphW += w2;
          for (int c = 0; c < D; ++c) af.xsphDV[c] = fmaR(-vij[c], w2, af.xsphDV[c]);
        }
      } else {
        // FP64, search row order: every accumulation fused
        for (int c = 0; c < D; ++c) a.DvDt.c[c] = fmaR(-rij[c], fsr, a.DvDt.c[c]);
        const Real mv = mj * vdr;
        a.DepsDt = fmaR(mv, pi, a.DepsDt);
        a.viscW = fmaR(mv, halfQi, a.viscW);
        a.effVP = fmaR(nb.mor * Qi, Wi, a.effVP);
        if (sameList) {
          a.rhoSum = fmaR(mj, Wi, a.rhoSum);
          a.norm = fmaR(mori, Wi, a.norm);
        }
        a.wNS += absR(gWi);
        for (int c = 0; c < D; ++c) {
          const Real u = vij[c] * wg;
          for (int b = 0; b < D; ++b) a.DvDx.c[c * D + b] = fmaR(-u, rij[b], a.DvDx.c[c * D + b]);
        }
        int s2 = 0;
        for (int c = 0; c < D; ++c) {
          const Real rw = rij[c] * wg, rf = rij[c] * f2;
          for (int b = c; b < D; ++b, ++s2) {
            a.M.c[c * D + b] = fmaR(-rw, rij[b], a.M.c[c * D + b]);
            a.m2nd.c[s2] = fmaR(rf, rij[b], a.m2nd.c[s2]);
          }
        }
        if (p.XSPH) {
          const Real w2 = Real(0.5) * fmaR(mori, Wi, nb.mor * Wj);
          a.xsphW += w2;
          for (int c = 0; c < D; ++c) a.xsphDV.c[c] = fmaR(-vij[c], w2, a.xsphDV.c[c]);
        }
      }
    };
    INbr<Real> nx;
    int32_t jn = cur.x;
    if (PREFETCH && len > 0) loadIso<D>(recN, n, jn, nx);
    for (int w = 0; w < len; ++w) {
      INbr<Real> nb;
      int32_t j;
      if constexpr (PREFETCH) {
        // the next record is loaded while this pair is evaluated
        nb = nx;
        j = jn;
        if (w + 1 < len) {
          jn = nextIndex(nb4, n, k, len, w + 1, cur, nxt);
          loadIso<D>(recN, n, jn, nx);
        }
      } else {
        // one record in flight per lane (fewer registers, more waves)
        j = w == 0 ? cur.x : nextIndex(nb4, n, k, len, w, cur, nxt);
        loadIso<D>(recN, n, j, nb);
      }
      pair(nb, j, w);
      if constexpr (F32) {
        if ((w & (kSumBlock - 1)) == kSumBlock - 1) af.flushInto(a);
      }
    }
    if constexpr (F32) af.flushInto(a);
    for (int c = 0; c < D; ++c)
      for (int b = 0; b < c; ++b) a.M.c[c * D + b] = a.M.c[b * D + c];
    a.maxVP = double(maxVP);
    // the node's own state for the epilogue, read again after the loop so it
    // holds no registers during it
    Self<D> me;
    for (int c = 0; c < D; ++c) me.v.c[c] = ri_[R::V + c];
    const double hinvD = ri_[RL<true>::HINV];
    for (int c = 0; c < NSym<D>::value; ++c) me.H.c[c] = 0.0;
    for (int c = 0; c < D; ++c) me.H.c[sdiag<D>(c)] = hinvD;
    me.Hdet = ri_[R::HDET]; me.m = ri_[R::M]; me.rho = ri_[R::RHO]; me.mor = ri_[R::MOR];
    epilogue<D>(k, len, me, a, t, p, o);
  }
  for (int sh = 32; sh > 0; sh >>= 1) pairs += __shfl_xor(pairs, sh);
  if ((threadIdx.x & 63) == 0 && pairs) atomicAdd(ncalc, (unsigned long long)pairs);
}

TabArgs tabArgs(const Sim& s) {
  return TabArgs{s.dTable.as<double>(), s.dTable32.as<float>(), s.table.n, s.table.step, s.table.extent,
                 s.dNperh.as<double>(), s.dWsum.as<double>()};
}

}  // namespace

double quantInv(const Sim& s) {
  // only the 3-D mixed record holds quantised positions; with no finite search
  // radius there is no pair to difference and any unit serves
  if (s.dim != 3 || !(s.rMax > 0.0) || !std::isfinite(s.rMax)) return 1.0;
  const int e = std::ilogb(s.rMax) + 1 - 30;     // rMax < 2^30 u
  SG_REQUIRE(e > -120 && e < 120, SPHGPU_ERR_UNSUPPORTED,
             "mixed precision: the search radius is outside the FP32 exponent range");
  const double q = std::ldexp(1.0, -e);
  // round(x / u) must fit a 64-bit integer; 2^61 leaves room for the motion
  // of one step beyond the positions the neighbour build saw
  SG_REQUIRE(s.posAbsMax * q < std::ldexp(1.0, 61), SPHGPU_ERR_UNSUPPORTED,
             "mixed precision: positions too far from the origin for the 3-D fixed-point record (|x| >= 2^61 u, "
             "u = largest search radius / 2^30)");
  return q;
}

PackIn packInputs(Sim& s) {
  return PackIn{s.perm.as<int32_t>(), s.pos.as<double>(), s.vel.as<double>(), s.mass.as<double>(),
                s.rho.as<double>(), s.P.as<double>(), s.cs.as<double>(), s.H.as<double>(), s.omega.as<double>(),
                s.hasMultipliers ? s.fCl.as<double>() : nullptr, s.hasMultipliers ? s.fCq.as<double>() : nullptr,
                s.hasMultipliers ? s.fshear.as<double>() : nullptr};
}

PackOut packOutputs(Sim& s, int sec) {
  const int64_t n = s.nSlots;
  const bool iso = s.isotropic;
  SG_REQUIRE(sec == 0 || iso, SPHGPU_ERR_STATE, "the pair kernels' records need H = h I");
  SG_REQUIRE((sec & 3) != 3, SPHGPU_ERR_STATE, "one pair-kernel record per pack");
  const int stride = iso ? RL<true>::STRIDE : RL<false>::STRIDE;
  s.rec.ensure(sizeof(double) * size_t(n) * stride);
  PackOut o{s.rec.as<double>(), nullptr, nullptr, n, s.nInt, nullptr, (sec & 1) ? quantInv(s) : 1.0};
  if (sec & 1) {
    s.recM.ensure(sizeof(float) * 16 * size_t(std::max<int64_t>(n, 1)));
    o.sec = s.recM.p;
  } else if (sec & 2) {
    s.recD.ensure(sizeof(double) * 12 * size_t(std::max<int64_t>(n, 1)));
    o.sec = s.recD.p;
  }
  if (sec & 4) {
    s.recS.ensure(sizeof(double) * DensRec::STRIDE * size_t(std::max<int64_t>(n, 1)));
    o.sec2 = s.recS.p;
  }
  if (sec & 8) {
    s.recSM.ensure(sizeof(float) * 8 * size_t(std::max<int64_t>(n, 1)));
    o.sec3 = s.recSM.p;
  }
  s.recSValid = (sec & 4) != 0;
  s.recSMValid = (sec & 8) != 0;
  s.pairRec = sec & 3;
  return o;
}

template <int D>
void launchPack(Sim& s, int sec, const int32_t* slots, int64_t count) {
  if (count <= 0) return;
  const PackOut o = packOutputs(s, sec);
  const PackIn in = packInputs(s);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nblocks(count)), dim3(kBlock), 0, s.stream, in, o, slots, count);
  };
  if (!s.isotropic) {
    go(k_pack<D, false, 0>);
  } else {
    switch (sec) {
      case 1: go(k_pack<D, true, 1>); break;
      case 2: go(k_pack<D, true, 2>); break;
      case 4: go(k_pack<D, true, 4>); break;
      case 5: go(k_pack<D, true, 5>); break;
      case 6: go(k_pack<D, true, 6>); break;
      case 8: go(k_pack<D, true, 8>); break;
      case 9: go(k_pack<D, true, 9>); break;
      default: go(k_pack<D, true, 0>); break;
    }
  }
  SG_CHECK_HIP(hipGetLastError());
}

// Records from the caller-order fields; sec bits 1 / 2 / 4 also write the
// mixed or FP64 isotropic pair kernel's record and the density sum's record
// in the same pass.
template <int D>
void packD(Sim& s, int sec = 0) {
  // With every slot live (no culled ghost), walk the caller rows so that the
  // field reads are contiguous and each slot's record is written whole; else
  // walk the slots, which also writes the dead ones.
  if (s.nSlots == s.nTot) launchPack<D>(s, sec, s.iperm.as<int32_t>(), s.nTot);
  else launchPack<D>(s, sec, nullptr, s.nSlots);
}

void packGhostRecords(Sim& s, int sec) {
  if (s.nTot == s.nInt) { (void)packOutputs(s, sec); return; }
  const int32_t* slots = s.iperm.as<int32_t>() + s.nInt;
  if (s.dim == 1) launchPack<1>(s, sec, slots, s.nTot - s.nInt);
  else if (s.dim == 2) launchPack<2>(s, sec, slots, s.nTot - s.nInt);
  else launchPack<3>(s, sec, slots, s.nTot - s.nInt);
}

int pairRecordSec(const Sim& s, const sphgpu_hydro_params& hp) {
  const bool tensile = hp.epsTensile != 0.0;
  const bool compat = hp.compatibleEnergyEvolution != 0;
  if (hp.precision == SPHGPU_PRECISION_MIXED) return 1;
  // search row order is held to round-off of the reference sums, not to
  // their bits: the collapsed isotropic algebra in FP64 (one NodeList)
  return (s.rowOrder == SPHGPU_ROW_ORDER_SEARCH && s.isotropic && !tensile && !compat && s.nLists <= 1) ? 2 : 0;
}

int densityRecordSec(const Sim& s) {
  if (!s.isotropic) return 0;
  return s.precision == SPHGPU_PRECISION_MIXED ? 8 : 4;
}

namespace {

Tab32 tab32(const Sim& s) { return Tab32{float(1.0 / s.table.step), float(s.table.extent), s.table.n}; }

template <int D>
void packMixed(Sim& s) {
  s.recM.ensure(sizeof(float) * 16 * size_t(std::max<int64_t>(s.nSlots, 1)));
  hipLaunchKernelGGL(k_pack_mixed<D>, dim3(nblocks(s.nSlots)), dim3(kBlock), 0, s.stream, s.rec.as<double>(), s.nSlots,
                     s.nInt, quantInv(s), s.recM.as<float4>());
  s.pairRec = 1;
  SG_CHECK_HIP(hipGetLastError());
}

template <int D>
void packFast64(Sim& s) {
  s.recD.ensure(sizeof(double) * 12 * size_t(std::max<int64_t>(s.nSlots, 1)));
  hipLaunchKernelGGL(k_pack_fast64<D>, dim3(nblocks(s.nSlots)), dim3(kBlock), 0, s.stream, s.rec.as<double>(),
                     s.nSlots, s.nInt, s.recD.as<double2>());
  s.pairRec = 2;
  SG_CHECK_HIP(hipGetLastError());
}

Tab64 tab64(const Sim& s) { return Tab64{1.0 / s.table.step, s.table.extent, s.table.n}; }

__global__ void k_slot_lists(const int32_t* __restrict__ perm, const double* __restrict__ seg, int64_t nSlots,
                             int32_t* __restrict__ out) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k < nSlots) {
    const int32_t i = perm[k];
    out[k] = i >= 0 ? int32_t(seg[i]) : -1;
  }
}

// Several NodeLists: the NodeList of every slot of the current build (for the
// collapsed kernels, whose records hold no caller index); null with one.
const int32_t* slotLists(Sim& s) {
  if (s.nLists <= 1) return nullptr;
  s.slotList.ensure(sizeof(int32_t) * size_t(std::max<int64_t>(s.nSlots, 1)));
  hipLaunchKernelGGL(k_slot_lists, dim3(nblocks(s.nSlots)), dim3(kBlock), 0, s.stream, s.perm.as<int32_t>(),
                     s.seg.as<double>(), s.nSlots, s.slotList.as<int32_t>());
  SG_CHECK_HIP(hipGetLastError());
  return s.slotList.as<int32_t>();
}

template <int D>
void sumDensityD(Sim& s) {
  const int64_t n = s.nSlots;
  TabArgs t = tabArgs(s);
  const bool lds = ldsTable(s);
  const size_t ldsBytes = lds ? sizeof(double) * 6 * size_t(s.table.n) : 0;
  if (!s.capturing) SG_CHECK_HIP(hipEventRecord(s.ev[2], s.stream));
  if (s.precision == SPHGPU_PRECISION_MIXED) {
    SG_REQUIRE(s.isotropic, SPHGPU_ERR_UNSUPPORTED, "mixed precision needs H = h I and no Q multipliers");
    if (!s.recSMValid) {
      s.recSM.ensure(sizeof(float) * 8 * size_t(std::max<int64_t>(n, 1)));
      hipLaunchKernelGGL(k_pack_dens_mixed<D>, dim3(nblocks(n)), dim3(kBlock), 0, s.stream, s.rec.as<double>(), n,
                         s.recSM.as<float4>());
      SG_CHECK_HIP(hipGetLastError());
      s.recSMValid = true;
    }
    const size_t lds32 = lds ? sizeof(float) * 8 * size_t(s.table.n) : 0;
    const double W0 = s.table.kernelValue(0.0, 1.0);
    const int32_t* lists = slotLists(s);
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(xcdGrid(nblocks(s.nInt))), dim3(kBlock), lds32, s.stream, s.rec.as<double>(),
                         s.recSM.as<float4>(), s.rowCount.as<int32_t>(), s.nbr.as<int32_t>(), n, s.nInt,
                         s.work.as<int32_t>(), t, tab32(s), W0, s.rho.as<double>(), lists, s.sumAllLists ? 1 : 0);
    };
    if (lds) go(k_sum_density_mixed<D, true>); else go(k_sum_density_mixed<D, false>);
    SG_CHECK_HIP(hipGetLastError());
    applyGhostDensity(s);
    if (!s.capturing) SG_CHECK_HIP(hipEventRecord(s.ev[3], s.stream));
    return;
  }
  // isotropic H: neighbours from the 48-B density records
  if (s.isotropic && !s.recSValid) {
    s.recS.ensure(sizeof(double) * DensRec::STRIDE * size_t(std::max<int64_t>(n, 1)));
    hipLaunchKernelGGL(k_pack_dens<D>, dim3(nblocks(n)), dim3(kBlock), 0, s.stream, s.rec.as<double>(), n,
                       s.recS.as<double2>());
    SG_CHECK_HIP(hipGetLastError());
    s.recSValid = true;
  }
  const double* nrec = s.isotropic ? s.recS.as<double>() : s.rec.as<double>();
  const double* listOf = s.nLists > 1 ? s.seg.as<double>() : nullptr;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(xcdGrid(nblocks(s.nInt))), dim3(kBlock), ldsBytes, s.stream, s.rec.as<double>(),
                       s.rowCount.as<int32_t>(), s.nbr.as<int32_t>(), n, s.nInt, s.work.as<int32_t>(), t,
                       s.rho.as<double>(), nrec, listOf, s.sumAllLists ? 1 : 0);
  };
  if (s.isotropic) { if (lds) go(k_sum_density<D, true, true>); else go(k_sum_density<D, true, false>); }
  else { if (lds) go(k_sum_density<D, false, true>); else go(k_sum_density<D, false, false>); }
  SG_CHECK_HIP(hipGetLastError());
  // SPHHydroBase::finalize then applies the boundaries to the mass density
  applyGhostDensity(s);
  if (!s.capturing) SG_CHECK_HIP(hipEventRecord(s.ev[3], s.stream));
}

// The collapsed isotropic pair kernel with BS-thread workgroups, its table in
// LDS or read through L1.  Tuning variant v: with a table that fits in LDS,
// BS = 64 << (v / 2) and the table in LDS when v is odd; otherwise BS = 64 << v.
template <typename Real> typename IsoTab<Real>::T isoTable(const Sim& s);
template <> Tab32 isoTable<float>(const Sim& s) { return tab32(s); }
template <> Tab64 isoTable<double>(const Sim& s) { return tab64(s); }
template <typename Real, int D, bool PF, int BS>
void isoLaunch(Sim& s, bool inLds, hipStream_t st, int64_t n, const TabArgs& t, const PassArgs& p, const OutArgs& o,
               unsigned long long* dncalc) {
  using Rec = typename IsoTab<Real>::Rec;
  const Rec* recN = reinterpret_cast<const Rec*>(std::is_same<Real, float>::value ? s.recM.p : s.recD.p);
  const size_t lb = inLds ? sizeof(typename IsoTab<Real>::Elem) * IsoTab<Real>::PER * size_t(s.table.n) : 0;
  const dim3 g(unsigned(xcdGrid(nblocks(s.nInt, BS)))), b(BS);
  if (inLds)
    hipLaunchKernelGGL((k_derivs_iso<Real, D, true, PF, BS>), g, b, lb, st, s.rec.as<double>(), recN,
                       s.rowCount.as<int32_t>(), s.nbr.as<int32_t>(), n, s.nInt, s.work.as<int32_t>(), t,
                       isoTable<Real>(s), p, o, dncalc);
  else
    hipLaunchKernelGGL((k_derivs_iso<Real, D, false, PF, BS>), g, b, 0, st, s.rec.as<double>(), recN,
                       s.rowCount.as<int32_t>(), s.nbr.as<int32_t>(), n, s.nInt, s.work.as<int32_t>(), t,
                       isoTable<Real>(s), p, o, dncalc);
}
template <typename Real, int D, bool PF>
void isoVariant(Sim& s, int v, bool fits, hipStream_t st, int64_t n, const TabArgs& t, const PassArgs& p,
                const OutArgs& o, unsigned long long* dncalc) {
  const int bsLog = fits ? v >> 1 : v;
  const bool inLds = fits && (v & 1);
  if constexpr (D == 3) {
    isoLaunch<Real, D, PF, 256>(s, inLds, st, n, t, p, o, dncalc);
  } else {
    switch (bsLog) {
      case 0: isoLaunch<Real, D, PF, 64>(s, inLds, st, n, t, p, o, dncalc); break;
      case 1: isoLaunch<Real, D, PF, 128>(s, inLds, st, n, t, p, o, dncalc); break;
      case 2: isoLaunch<Real, D, PF, 256>(s, inLds, st, n, t, p, o, dncalc); break;
      default: isoLaunch<Real, D, PF, 512>(s, inLds, st, n, t, p, o, dncalc); break;
    }
  }
  SG_CHECK_HIP(hipGetLastError());
}

template <int D>
void evalDerivsD(Sim& s, const sphgpu_hydro_params& hp, bool updateOnly) {
  const int64_t n = s.nSlots;
  hipStream_t st = s.stream;
  const int total = derivTotalDoubles(s.dim);
  s.out.ensure(sizeof(double) * size_t(n) * total);
  s.outStride = n;
  s.reduceTmp.ensure(sizeof(double) * 16);
  unsigned long long* dncalc = reinterpret_cast<unsigned long long*>(s.reduceTmp.as<double>() + 8);

  // the pair kernel decides the record it reads
  const bool tensile = hp.epsTensile != 0.0;
  const bool compat = hp.compatibleEnergyEvolution != 0;
  const bool mixed = hp.precision == SPHGPU_PRECISION_MIXED;
  if (mixed)
    SG_REQUIRE(s.isotropic && !tensile, SPHGPU_ERR_UNSUPPORTED,
               "mixed precision needs H = h I, no Q multipliers and epsTensile = 0");
  const int sec = pairRecordSec(s, hp);
  const bool fast64 = sec == 2;

  if (!s.capturing) SG_CHECK_HIP(hipEventRecord(s.ev[4], st));
  if (!s.recValid) {
    packD<D>(s, sec);
    s.recValid = true;
  } else if (sec == 1 && s.pairRec != 1) {
    packMixed<D>(s);
  } else if (sec == 2 && s.pairRec != 2) {
    packFast64<D>(s);
  }
  {
    // a kernel, not a memset node, when the step is captured as a graph
    ZeroSet z{};
    z.add(reinterpret_cast<int32_t*>(dncalc), 2);
    zeroMany(z, st);
  }
  if (!s.capturing) SG_CHECK_HIP(hipEventRecord(s.ev[5], st));

  TabArgs t = tabArgs(s);
  const bool lds = ldsTable(s);
  const size_t ldsBytes = lds ? sizeof(double) * 6 * size_t(s.table.n) : 0;
  PassArgs p{};
  p.Cl = hp.Cl; p.Cq = hp.Cq; p.eps2 = hp.eps2; p.epsTensile = hp.epsTensile; p.nPerh = hp.nPerh;
  p.hmin = hp.hmin; p.hmax = hp.hmax;
  p.W0 = s.table.kernelValue(0.0, 1.0);
  p.WnPerh = s.table.kernelValue(1.0 / hp.nPerh, 1.0);
  p.XSPH = hp.XSPH; p.correctVelocityGradient = hp.correctVelocityGradient; p.evolveTotalEnergy = hp.evolveTotalEnergy;
  p.linearInExpansion = hp.linearInExpansion; p.quadraticInExpansion = hp.quadraticInExpansion;
  p.qInv = mixed ? quantInv(s) : 1.0;
  p.qUnit = float(1.0 / p.qInv);
  p.listOf = s.nLists > 1 ? s.seg.as<double>() : nullptr;
  p.slotList = mixed ? slotLists(s) : nullptr;
  OutArgs o{};
  o.base = s.out.as<double>();
  o.n = n;
  o.updateOnly = updateOnly ? 1 : 0;
  for (int f = 0; f < SPHGPU_NUM_DERIV_FIELDS; ++f) o.off[f] = derivOffsetDoubles(s.dim, f);
  if (compat) {
    // pairAccelerations: the own-frame pair term of every row entry
    o.plane = n * int64_t(s.ellCap);
    s.pairDv.ensure(sizeof(double) * size_t(D) * size_t(std::max<int64_t>(o.plane, 1)));
    o.pairDv = s.pairDv.as<double>();
  }
  // Mixed: one record in flight per lane keeps the kernel at 167 VGPRs, three
  // wavefronts per SIMD (2.11 ms at 128^3 against 2.53 ms with the next
  // record prefetched at 179 VGPRs, two wavefronts).  FP64: the prefetch wins
  // (3.44 ms at 194 VGPRs against 3.51 ms at 168).
  if (mixed || fast64) {
    // 1-D and 2-D (ensembles, serving): the workgroup size and the table's
    // place (LDS or read through L1) are autotuned per workload class.  3-D
    // runs 256-thread workgroups with the table in LDS when it fits.
    const bool fits = s.table.n <= kLdsTablePoints;
    const int nv = fits ? 8 : 4;
    auto run = [&](int v) {
      if (mixed) isoVariant<float, D, false>(s, v, fits, st, n, t, p, o, dncalc);
      else isoVariant<double, D, true>(s, v, fits, st, n, t, p, o, dncalc);
    };
    const int def = fits ? (4 + (lds ? 1 : 0)) : 2;      // 256 threads, the row-length rule for the table
    int v = def;
    if (D < 3) {
      bool timed = false;
      const double perNode = s.nInt > 0 ? double(s.nEntries) / double(s.nInt) : 0.0;
      v = tunedVariant(s, tuneKey(kTunePairIso, D, mixed ? 1 : 0, perNode, s.nInt), nv, def, run, &timed);
      if (timed) {
        // the trials added to the pair count
        ZeroSet z{};
        z.add(reinterpret_cast<int32_t*>(dncalc), 2);
        zeroMany(z, st);
      }
    }
    run(v);
  } else {
    auto launch = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(xcdGrid(nblocks(s.nInt))), dim3(kBlock), ldsBytes, st, s.rec.as<double>(),
                         s.rowCount.as<int32_t>(), s.nbr.as<int32_t>(), n, s.nInt, s.work.as<int32_t>(), t, p, o,
                         dncalc);
    };
    if (lds) {
      if (s.isotropic) { if (tensile) launch(k_derivs_fp64<D, true, true, true>); else launch(k_derivs_fp64<D, true, false, true>); }
      else { if (tensile) launch(k_derivs_fp64<D, false, true, true>); else launch(k_derivs_fp64<D, false, false, true>); }
    } else {
      if (s.isotropic) { if (tensile) launch(k_derivs_fp64<D, true, true, false>); else launch(k_derivs_fp64<D, true, false, false>); }
      else { if (tensile) launch(k_derivs_fp64<D, false, true, false>); else launch(k_derivs_fp64<D, false, false, false>); }
    }
  }
  SG_CHECK_HIP(hipGetLastError());
  if (!s.capturing) SG_CHECK_HIP(hipEventRecord(s.ev[6], st));
  SG_CHECK_HIP(hipMemcpyAsync(s.ncalcHost, dncalc, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
}

}  // namespace

template <int D>
void omegaGradhD(Sim& s) {
  const int64_t n = s.nSlots;
  TabArgs t = tabArgs(s);
  const bool lds = ldsTable(s);
  const size_t ldsBytes = lds ? sizeof(double) * 6 * size_t(s.table.n) : 0;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(xcdGrid(nblocks(s.nInt))), dim3(kBlock), ldsBytes, s.stream, s.rec.as<double>(),
                       s.rowCount.as<int32_t>(), s.nbr.as<int32_t>(), n, s.nInt, s.work.as<int32_t>(), t,
                       s.omega.as<double>());
  };
  if (s.isotropic) { if (lds) go(k_omega<D, true, true>); else go(k_omega<D, true, false>); }
  else { if (lds) go(k_omega<D, false, true>); else go(k_omega<D, false, false>); }
  SG_CHECK_HIP(hipGetLastError());
}

void omegaGradh(Sim& s) {
  SG_REQUIRE(s.neighborsValid, SPHGPU_ERR_STATE, "the grad-h correction needs a neighbour build");
  SG_REQUIRE(s.table.n > 0, SPHGPU_ERR_STATE, "no kernel table installed");
  if (!s.recValid) packRecords(s);
  if (s.dim == 1) omegaGradhD<1>(s);
  else if (s.dim == 2) omegaGradhD<2>(s);
  else omegaGradhD<3>(s);
  // SPHHydroBase::initialize applies the boundaries to omega
  applyGhostScalar(s, s.omega);
  s.recValid = false;         // records hold safeInv(omega)
  s.derivsValid = false;
}

// One iterateIdealH pass (src/Utilities/iterateIdealH.cc:139-230) with the
// SPH smoothing scale: one lane per internal node; a node whose last change is
// within tol keeps its H, the others take zeroth = rootnu(sum over the row of
// |W.gradValue(|H_i x_ij|, 1)|) in row order, H1 = idealSmoothingScale, and the
// change max|phi - 1| over the eigenvalues c/lambda(H_i) of H1^1/2 H_i^-1 H1^1/2
// (H1 = c I).  Node i reads only its own H (neighbour positions come from
// the records), so H is replaced in place.
template <int D, bool ISO, bool LDS>
__global__ void __launch_bounds__(kBlock) k_ideal_h(const double* __restrict__ rec, const int32_t* __restrict__ rowCount,
                                                    const int32_t* __restrict__ nbr, int64_t n, int64_t nInt,
                                                    const int32_t* __restrict__ work, const int32_t* __restrict__ perm,
                                                    TabArgs t, double nPerh, double hmin, double hmax, double tol,
                                                    double* __restrict__ H, double* __restrict__ deltaH,
                                                    unsigned long long* __restrict__ maxBits,
                                                    const double* __restrict__ listOf) {
  using R = RL<ISO>;
  constexpr int NS = NSym<D>::value;
  extern __shared__ double lds[];
  const double* tab = stageTable<LDS>(lds, t);
  const int64_t tw = xcdBlock((nInt + kBlock - 1) / kBlock) * kBlock + threadIdx.x;
  double d = 0.0;
  if (tw < nInt) {
    const int64_t k = work[tw];
    const int64_t i = perm[k];
    if (deltaH[i] > tol) {
      S<D> Hi;
      for (int c = 0; c < NS; ++c) Hi.c[c] = H[i * NS + c];
      const double* ri_ = rec + k * R::STRIDE;
      V<D> xi;
      for (int c = 0; c < D; ++c) xi.c[c] = ri_[R::X + c];
      // several NodeLists: a neighbour of another NodeList counts with
      // m_j rho_i / (m_i rho_j) (iterateIdealH.cc:183-194)
      const double li = listOf ? listOf[i] : 0.0, mi = ri_[R::M], rhoi = ri_[R::RHO];
      double zeroth = 0.0;
      const int len = rowCount[k];
      for (int w = 0; w < len; ++w) {
        const double* rj = rec + int64_t(nbr[ellIndex(w, n, k)]) * R::STRIDE;
        V<D> xij;
        for (int c = 0; c < D; ++c) xij.c[c] = xi.c[c] - rj[R::X + c];
        const double etai = vmag(sdot(Hi, xij));
        double Wv, gW;
        lookupWG(tab, t, etai, 1.0, Wv, gW);
        double fw = 1.0;
        if (listOf && listOf[int64_t(rj[R::ORIG])] != li) fw = rj[R::M] * rhoi / (mi * rj[R::RHO]);
        zeroth += fw * dabs(gW);
      }
      zeroth = rootnu<D>(zeroth);
      const double c1 = idealHinv(t, zeroth, Hi.c[0], nPerh, hmin, hmax);
      const double phimin = c1 / maxEigen<D>(Hi.c), phimax = c1 / minEigen<D>(Hi.c);
      d = dmax(dabs(phimin - 1.0), dabs(phimax - 1.0));
      deltaH[i] = d;
      int q = 0;
      for (int a = 0; a < D; ++a)
        for (int b = a; b < D; ++b, ++q) H[i * NS + q] = (a == b) ? c1 : 0.0;
    }
  }
  for (int sh = 32; sh > 0; sh >>= 1) d = dmax(d, __shfl_xor(d, sh));
  if ((threadIdx.x & 63) == 0 && d > 0.0) atomicMax(maxBits, (unsigned long long)__double_as_longlong(d));
}

template <int D>
double idealHPassD(Sim& s, const sphgpu_hydro_params& hp, double tol, double* deltaH, unsigned long long* bits) {
  const int64_t n = s.nSlots;
  TabArgs t = tabArgs(s);
  const bool lds = ldsTable(s);
  const size_t ldsBytes = lds ? sizeof(double) * 6 * size_t(s.table.n) : 0;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(xcdGrid(nblocks(s.nInt))), dim3(kBlock), ldsBytes, s.stream, s.rec.as<double>(),
                       s.rowCount.as<int32_t>(), s.nbr.as<int32_t>(), n, s.nInt, s.work.as<int32_t>(),
                       s.perm.as<int32_t>(), t, hp.nPerh, hp.hmin, hp.hmax, tol, s.H.as<double>(), deltaH, bits,
                       s.nLists > 1 ? s.seg.as<double>() : nullptr);
  };
  if (s.isotropic) { if (lds) go(k_ideal_h<D, true, true>); else go(k_ideal_h<D, true, false>); }
  else { if (lds) go(k_ideal_h<D, false, true>); else go(k_ideal_h<D, false, false>); }
  SG_CHECK_HIP(hipGetLastError());
  unsigned long long hb = 0;
  SG_CHECK_HIP(hipMemcpyAsync(&hb, bits, sizeof(hb), hipMemcpyDeviceToHost, s.stream));
  SG_CHECK_HIP(hipStreamSynchronize(s.stream));
  double m;
  std::memcpy(&m, &hb, sizeof(m));
  return m;
}

double idealHPass(Sim& s, const sphgpu_hydro_params& hp, double tol, double* deltaH, unsigned long long* bits) {
  SG_REQUIRE(s.neighborsValid, SPHGPU_ERR_STATE, "the ideal H pass needs a neighbour build");
  SG_REQUIRE(s.table.n > 0, SPHGPU_ERR_STATE, "no kernel table installed");
  if (!s.recValid) packRecords(s);
  double m;
  if (s.dim == 1) m = idealHPassD<1>(s, hp, tol, deltaH, bits);
  else if (s.dim == 2) m = idealHPassD<2>(s, hp, tol, deltaH, bits);
  else m = idealHPassD<3>(s, hp, tol, deltaH, bits);
  s.recValid = false;
  s.derivsValid = false;
  return m;
}

void packRecords(Sim& s, int sec) {
  if (s.dim == 1) packD<1>(s, sec);
  else if (s.dim == 2) packD<2>(s, sec);
  else packD<3>(s, sec);
  s.recValid = true;
}

void sumDensity(Sim& s) {
  SG_REQUIRE(s.neighborsValid, SPHGPU_ERR_STATE, "sum density needs a neighbour build");
  SG_REQUIRE(s.table.n > 0, SPHGPU_ERR_STATE, "no kernel table installed");
  if (!s.recValid) {
    if (s.isotropic) {
      // the records and the density kernel's own record in one pass
      const int sec = densityRecordSec(s);
      if (s.dim == 1) packD<1>(s, sec);
      else if (s.dim == 2) packD<2>(s, sec);
      else packD<3>(s, sec);
      s.recValid = true;
    } else {
      packRecords(s);
    }
  }
  if (s.dim == 1) sumDensityD<1>(s);
  else if (s.dim == 2) sumDensityD<2>(s);
  else sumDensityD<3>(s);
  s.pairRec = 0;      // the density changed rho in the records
  if (!s.capturing) s.timingPending |= 2u;
  s.derivsValid = false;
}

void evalDerivs(Sim& s, const sphgpu_hydro_params& p, bool updateOnly) {
  SG_REQUIRE(s.neighborsValid, SPHGPU_ERR_STATE, "evaluateDerivatives needs a neighbour build");
  SG_REQUIRE(s.table.n > 0, SPHGPU_ERR_STATE, "no kernel table installed");
  SG_REQUIRE(!(p.compatibleEnergyEvolution && p.evolveTotalEnergy), SPHGPU_ERR_ARG,
             "compatibleEnergyEvolution and evolveTotalEnergy exclude each other");
  SG_REQUIRE(p.precision == SPHGPU_PRECISION_FP64 || p.precision == SPHGPU_PRECISION_MIXED, SPHGPU_ERR_ARG,
             "unknown precision");
  SG_REQUIRE(p.nPerh > 0.0, SPHGPU_ERR_ARG, "nPerh must be positive");
  s.derivsValid = false;
  if (s.dim == 1) evalDerivsD<1>(s, p, updateOnly);
  else if (s.dim == 2) evalDerivsD<2>(s, p, updateOnly);
  else evalDerivsD<3>(s, p, updateOnly);
  if (!s.capturing) s.timingPending |= 4u;
  s.derivsValid = true;
  s.pairsValid = true;
}

int derivComponents(int dim, int f) {
  const int ns = nsymOf(dim);
  switch (f) {
    case SPHGPU_D_DXDT: case SPHGPU_D_DVDT: case SPHGPU_D_XSPH_DELTA_V: return dim;
    case SPHGPU_D_DVDX: case SPHGPU_D_LOCAL_DVDX: case SPHGPU_D_M: case SPHGPU_D_LOCAL_M: return dim * dim;
    case SPHGPU_D_DHDT: case SPHGPU_D_HIDEAL: case SPHGPU_D_MASS_SECOND_MOMENT: return ns;
    default: return 1;
  }
}
int derivOffsetDoubles(int dim, int f) {
  int o = 0;
  for (int q = 0; q < f; ++q) o += derivComponents(dim, q);
  return o;
}
int derivTotalDoubles(int dim) { return derivOffsetDoubles(dim, SPHGPU_NUM_DERIV_FIELDS); }

}  // namespace sphgpu
