// Device sum density and evaluateDerivatives.
//
// Both passes run in gather form: one lane per internal node (nodes in the
// neighbour build's sorted order), walking its full row in ascending node index.
// The reference (src/SPH/SPHHydroBase.cc:575-928,
// src/SPH/computeSPHSumMassDensity.cc:60-103) evaluates each unordered pair once
// in the frame of its owner (the lower index, ConnectivityMapInline.hh:182-184)
// and adds to both ends, visiting owners in ascending order; so every node
// receives its pair terms in ascending neighbour order.  Walking the row in that
// order, in the node's own frame, reproduces the reference's FP64 sums exactly:
//   * quantities odd in r_ij (eta, unit vectors, grad W, v_ij, Q accelerations,
//     the acceleration increment) only flip sign with the frame, exactly;
//   * products of two odd quantities (mu, v.gradW, dyads, r r) are unchanged;
//   * the two terms that depend on who owns the pair are kept: the Q
//     accelerations enter the acceleration sum owner first, and the tensile
//     factor uses the owner's W.
// Sum density adds the self term where the reference does: after the
// contributions of lower-index neighbours, before those of higher ones.
#include "internal.hh"
#include "dgeom.hh"

namespace sphgpu {

namespace {

constexpr int kBlock = 256;
inline int nblocks(int64_t n, int bs = kBlock) { return int(std::max<int64_t>(1, (n + bs - 1) / bs)); }

// FP64 node record in sorted order: everything one side of a pair needs.
template <int D> struct Rec {
  static constexpr int NS = NSym<D>::value;
  static constexpr int X = 0, V = D, H = 2 * D, HDET = 2 * D + NS, M = HDET + 1, RHO = M + 1, CS = RHO + 1, P = CS + 1,
                       SOM = P + 1, PRHO = SOM + 1, MOR = PRHO + 1, FCL = MOR + 1, FCQ = FCL + 1, FSH = FCQ + 1;
  static constexpr int N = ((FSH + 1) + 1) & ~1;   // even -> 16-byte rows
};

template <int D>
__global__ void k_pack(const int32_t* __restrict__ perm, int64_t n, const double* __restrict__ pos,
                       const double* __restrict__ vel, const double* __restrict__ mass, const double* __restrict__ rho,
                       const double* __restrict__ P, const double* __restrict__ cs, const double* __restrict__ H,
                       const double* __restrict__ omega, const double* __restrict__ fCl, const double* __restrict__ fCq,
                       const double* __restrict__ fsh, double* __restrict__ rec) {
  using R = Rec<D>;
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k >= n) return;
  const int64_t i = perm[k];
  double* r = rec + k * R::N;
  for (int c = 0; c < D; ++c) { r[R::X + c] = pos[i * D + c]; r[R::V + c] = vel[i * D + c]; }
  S<D> h;
  for (int c = 0; c < R::NS; ++c) { h.c[c] = H[i * R::NS + c]; r[R::H + c] = h.c[c]; }
  const double rh = rho[i], m = mass[i], p = P[i];
  const double som = safeInv(omega[i], 1.0e-30);
  r[R::HDET] = sdet(h);
  r[R::M] = m;
  r[R::RHO] = rh;
  r[R::CS] = cs[i];
  r[R::P] = p;
  r[R::SOM] = som;
  r[R::PRHO] = som * p / (rh * rh);
  r[R::MOR] = m / rh;
  r[R::FCL] = fCl ? fCl[i] : 1.0;
  r[R::FCQ] = fCq ? fCq[i] : 1.0;
  r[R::FSH] = fsh ? fsh[i] : 1.0;
  for (int c = R::FSH + 1; c < R::N; ++c) r[c] = 0.0;
}

// Table in LDS: per point [Ak Bk Ck Ag Bg Cg].
struct TabArgs {
  const double* g;     // global copy
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

__device__ __forceinline__ const double* stageTable(double* lds, const TabArgs& t, bool useLds) {
  if (useLds) {
    for (int q = threadIdx.x; q < 6 * t.n; q += blockDim.x) lds[q] = t.g[q];
    __syncthreads();
    return lds;
  }
  return t.g;
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

// ---------------------------------------------------------------------------
// computeSPHSumMassDensity
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(kBlock) k_sum_density(const double* __restrict__ xs, const double* __restrict__ Hs,
                                                        const double* __restrict__ mass, const int32_t* __restrict__ perm,
                                                        const int64_t* __restrict__ off, const int32_t* __restrict__ nbr,
                                                        int64_t n, int64_t nInt, TabArgs t, bool useLds,
                                                        double* __restrict__ rhoOut) {
  extern __shared__ double lds[];
  const double* tab = stageTable(lds, t, useLds);
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k >= n) return;
  const int32_t oi = perm[k];
  if (oi >= nInt) return;
  constexpr int NS = NSym<D>::value;
  V<D> ri; S<D> Hi;
  for (int c = 0; c < D; ++c) ri.c[c] = xs[k * D + c];
  for (int c = 0; c < NS; ++c) Hi.c[c] = Hs[k * NS + c];
  const double self = mass[oi] * lookupW(tab, t, 0.0, sdet(Hi));
  double acc = 0.0;
  bool selfDone = false;
  const int64_t e = off[k + 1];
  for (int64_t q = off[k]; q < e; ++q) {
    const int32_t j = nbr[q];
    const int32_t oj = perm[j];
    if (!selfDone && oj > oi) { acc += self; selfDone = true; }
    V<D> rij; S<D> Hj;
    for (int c = 0; c < D; ++c) rij.c[c] = ri.c[c] - xs[int64_t(j) * D + c];
    for (int c = 0; c < NS; ++c) Hj.c[c] = Hs[int64_t(j) * NS + c];
    const double etaj = vmag(sdot(Hj, rij));
    acc += mass[oj] * lookupW(tab, t, etaj, sdet(Hj));
  }
  if (!selfDone) acc += self;
  rhoOut[oi] = acc;
}

// ---------------------------------------------------------------------------
// evaluateDerivatives, FP64
// ---------------------------------------------------------------------------
struct PassArgs {
  double Cl, Cq, eps2, epsTensile, nPerh, hmin, hmax;
  double W0, WnPerh;
  int XSPH, correctVelocityGradient, evolveTotalEnergy, linearInExpansion, quadraticInExpansion;
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
};

template <int D>
__device__ __forceinline__ double* outRow(const OutArgs& o, int field, int64_t k, int comps) {
  return o.base + int64_t(o.off[field]) * o.n + k * comps;
}

template <int D, bool TENSILE, bool MULT>
__device__ __forceinline__ unsigned derivsNodeFP64(int64_t k, int32_t oi, const double* __restrict__ rec,
                                                   const int32_t* __restrict__ perm, const int64_t* __restrict__ off,
                                                   const int32_t* __restrict__ nbr, const double* __restrict__ tab,
                                                   const TabArgs& t, const PassArgs& p, const OutArgs& o) {
  using R = Rec<D>;
  constexpr int NS = NSym<D>::value;
  const double tiny = 1.0e-30;

  const double* ri_ = rec + k * R::N;
  V<D> ri, vi; S<D> Hi;
  for (int c = 0; c < D; ++c) { ri.c[c] = ri_[R::X + c]; vi.c[c] = ri_[R::V + c]; }
  for (int c = 0; c < NS; ++c) Hi.c[c] = ri_[R::H + c];
  const double Hdeti = ri_[R::HDET], mi = ri_[R::M], rhoi = ri_[R::RHO], ci = ri_[R::CS], Pi = ri_[R::P];
  const double somi = ri_[R::SOM], Prhoi0 = ri_[R::PRHO], mori = ri_[R::MOR];
  const double fcli = ri_[R::FCL], fcqi = ri_[R::FCQ], fshi = ri_[R::FSH];

  double rhoSum = 0.0, norm = 0.0, DepsDt = 0.0, maxVP = 0.0, effVP = 0.0, viscW = 0.0, xsphW = 0.0, wNS = 0.0;
  V<D> DvDt = vzero<D>(), xsphDV = vzero<D>();
  T<D> DvDx = tzero<D>(), M = tzero<D>();
  S<D> m2nd = szero<D>();
  unsigned pairs = 0;

  const int64_t e = off[k + 1];
  for (int64_t q = off[k]; q < e; ++q) {
    const int32_t j = nbr[q];
    const int32_t oj = perm[j];
    const bool ownerI = oj > oi;
    pairs += ownerI ? 1u : 0u;
    const double* rj_ = rec + int64_t(j) * R::N;
    V<D> rij, vij; S<D> Hj;
    for (int c = 0; c < D; ++c) { rij.c[c] = ri.c[c] - rj_[R::X + c]; vij.c[c] = vi.c[c] - rj_[R::V + c]; }
    for (int c = 0; c < NS; ++c) Hj.c[c] = rj_[R::H + c];
    const double Hdetj = rj_[R::HDET], mj = rj_[R::M], rhoj = rj_[R::RHO], cj = rj_[R::CS];
    const double morj = rj_[R::MOR];

    const V<D> etai = sdot(Hi, rij), etaj = sdot(Hj, rij);
    const double eta2i = vmag2(etai), eta2j = vmag2(etaj);
    const double etaMagi = (D == 1) ? fabs(etai.c[0]) : __builtin_sqrt(eta2i);
    const double etaMagj = (D == 1) ? fabs(etaj.c[0]) : __builtin_sqrt(eta2j);
    const V<D> Hetai = sdot(Hi, vunit(etai, etaMagi));
    const V<D> Hetaj = sdot(Hj, vunit(etaj, etaMagj));
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
    rhoSum += mj * Wi;
    norm += mori * Wi;

    // Monaghan-Gingold Q (MonaghanGingoldViscosity.cc:83-132)
    double Cl = p.Cl, Cq = p.Cq;
    if (MULT) {
      const double fsh = dmax(fshi, rj_[R::FSH]);
      Cl *= 0.5 * (fcli + rj_[R::FCL]) * fsh;
      Cq *= 0.5 * (fcqi + rj_[R::FCQ]) * fsh;
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

    double Prhoi = Prhoi0, Prhoj = rj_[R::PRHO];
    if (TENSILE) {
      const double Wo = ownerI ? Wi : Wj, Hdeto = ownerI ? Hdeti : Hdetj;
      const double x = Wo / (Hdeto * p.WnPerh);
      const double fij = p.epsTensile * (x * x * x * x);
      const double Pj = rj_[R::P];
      Prhoi = somi * (Pi + fij * (Pi < 0.0 ? -Pi : 0.0)) / (rhoi * rhoi);
      Prhoj = rj_[R::SOM] * (Pj + fij * (Pj < 0.0 ? -Pj : 0.0)) / (rhoj * rhoj);
    }
    for (int c = 0; c < D; ++c) {
      const double qf = ownerI ? Qacci.c[c] : Qaccj.c[c];
      const double qs = ownerI ? Qaccj.c[c] : Qacci.c[c];
      const double dv = gradWi.c[c] * Prhoi + gradWj.c[c] * Prhoj + qf + qs;
      DvDt.c[c] += dv * (-mj);
    }
    DepsDt += mj * (Prhoi * vdot(vij, gradWi) + workQi);
    for (int a = 0; a < D; ++a)
      for (int b = 0; b < D; ++b) {
        DvDx.c[a * D + b] -= (vij.c[a] * gradWi.c[b]) * mj;
        M.c[a * D + b] -= (rij.c[a] * gradWi.c[b]) * mj;
      }
    if (p.XSPH) {
      const double w = 0.5 * (mori * Wi + morj * Wj);
      xsphW += w;
      for (int c = 0; c < D; ++c) xsphDV.c[c] -= vij.c[c] * w;
    }
  }

  // per-node epilogue (SPHHydroBase.cc:834-902)
  const int64_t numNeighbors = e - off[k];
  rhoSum += mi * p.W0 * Hdeti;
  norm += mori * p.W0 * Hdeti;
  T<D> Mout = M;
  if (p.correctVelocityGradient && fabs(tdet(M)) > 1.0e-10 && double(numNeighbors) > pownu2<D>()) {
    Mout = tinv(M);
    DvDx = tdot(DvDx, Mout);
  } else {
    tdivEq(DvDx, rhoi);
  }
  const double trace = ttrace(DvDx);
  const double DrhoDt = -rhoi * trace;
  if (p.evolveTotalEnergy) DepsDt = mi * (vdot(vi, DvDt) + DepsDt);
  const double wNSo = rootnu<D>(dmax(0.0, wNS / Hdeti));
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
  double cur;
  if (fuzzyEqual(wNSo, 0.0)) cur = 0.5 * p.nPerh;
  else cur = equivalentNperh(t, wNSo);
  const double s = dmin(4.0, dmax(0.25, p.nPerh / (cur + 1.0e-30)));
  const double a = (s < 1.0) ? 0.4 * (1.0 + s * s) : 0.4 * (1.0 + 1.0 / (s * s * s));
  const double hi0 = 1.0 / Hi.c[0];
  const double hi1 = dmin(p.hmax, dmax(p.hmin, hi0 * (1.0 - a + a * s)));
  const double hinv = 1.0 / hi1;

  // outputs
  *outRow<D>(o, SPHGPU_D_RHO_SUM, k, 1) = rhoSum;
  *outRow<D>(o, SPHGPU_D_NORMALIZATION, k, 1) = norm;
  *outRow<D>(o, SPHGPU_D_DRHODT, k, 1) = DrhoDt;
  *outRow<D>(o, SPHGPU_D_DEPSDT, k, 1) = DepsDt;
  *outRow<D>(o, SPHGPU_D_MAX_VISCOUS_PRESSURE, k, 1) = maxVP;
  *outRow<D>(o, SPHGPU_D_EFF_VISCOUS_PRESSURE, k, 1) = effVP;
  *outRow<D>(o, SPHGPU_D_VISCOUS_WORK, k, 1) = viscW;
  *outRow<D>(o, SPHGPU_D_XSPH_WEIGHT_SUM, k, 1) = xsphW;
  *outRow<D>(o, SPHGPU_D_WEIGHTED_NEIGHBOR_SUM, k, 1) = wNSo;
  {
    double* dx = outRow<D>(o, SPHGPU_D_DXDT, k, D);
    double* dv = outRow<D>(o, SPHGPU_D_DVDT, k, D);
    double* xv = outRow<D>(o, SPHGPU_D_XSPH_DELTA_V, k, D);
    for (int c = 0; c < D; ++c) { dx[c] = DxDt.c[c]; dv[c] = DvDt.c[c]; xv[c] = xsphDV.c[c]; }
  }
  {
    double* g = outRow<D>(o, SPHGPU_D_DVDX, k, D * D);
    double* gl = outRow<D>(o, SPHGPU_D_LOCAL_DVDX, k, D * D);
    double* m = outRow<D>(o, SPHGPU_D_M, k, D * D);
    double* ml = outRow<D>(o, SPHGPU_D_LOCAL_M, k, D * D);
    for (int c = 0; c < D * D; ++c) { g[c] = DvDx.c[c]; gl[c] = DvDx.c[c]; m[c] = Mout.c[c]; ml[c] = Mout.c[c]; }
  }
  {
    double* dh = outRow<D>(o, SPHGPU_D_DHDT, k, NS);
    double* hid = outRow<D>(o, SPHGPU_D_HIDEAL, k, NS);
    double* m2 = outRow<D>(o, SPHGPU_D_MASS_SECOND_MOMENT, k, NS);
    int s2 = 0;
    for (int a2 = 0; a2 < D; ++a2)
      for (int b2 = a2; b2 < D; ++b2, ++s2) {
        dh[s2] = DHDt.c[s2];
        hid[s2] = (a2 == b2 ? 1.0 : 0.0) * hinv;
        m2[s2] = m2nd.c[s2];
      }
  }
  return pairs;
}

template <int D, bool TENSILE, bool MULT>
__global__ void __launch_bounds__(kBlock) k_derivs_fp64(const double* __restrict__ rec, const int32_t* __restrict__ perm,
                                                        const int64_t* __restrict__ off, const int32_t* __restrict__ nbr,
                                                        int64_t n, int64_t nInt, TabArgs t, bool useLds, PassArgs p,
                                                        OutArgs o, unsigned long long* __restrict__ ncalc) {
  extern __shared__ double lds[];
  const double* tab = stageTable(lds, t, useLds);
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int32_t oi = k < n ? perm[k] : 0;
  unsigned pairs = 0;
  if (k < n && oi < nInt) pairs = derivsNodeFP64<D, TENSILE, MULT>(k, oi, rec, perm, off, nbr, tab, t, p, o);
  for (int sh = 32; sh > 0; sh >>= 1) pairs += __shfl_xor(pairs, sh);
  if ((threadIdx.x & 63) == 0 && pairs) atomicAdd(ncalc, (unsigned long long)pairs);
}

template <int D>
void sumDensityD(Sim& s) {
  const int64_t n = s.nTot;
  TabArgs t{s.dTable.as<double>(), s.table.n, s.table.step, s.table.extent, s.dNperh.as<double>(), s.dWsum.as<double>()};
  const size_t ldsBytes = sizeof(double) * 6 * size_t(s.table.n);
  const bool useLds = ldsBytes <= 64 * 1024;
  SG_CHECK_HIP(hipEventRecord(s.ev[2], s.stream));
  hipLaunchKernelGGL((k_sum_density<D>), dim3(nblocks(n)), dim3(kBlock), useLds ? ldsBytes : 0, s.stream,
                     s.xs.as<double>(), s.Hs.as<double>(), s.mass.as<double>(), s.perm.as<int32_t>(),
                     s.offsets.as<int64_t>(), s.nbr.as<int32_t>(), n, s.nInt, t, useLds, s.rho.as<double>());
  SG_CHECK_HIP(hipGetLastError());
  SG_CHECK_HIP(hipEventRecord(s.ev[3], s.stream));
}

template <int D>
void evalDerivsD(Sim& s, const sphgpu_hydro_params& hp) {
  using R = Rec<D>;
  const int64_t n = s.nTot;
  hipStream_t st = s.stream;
  s.rec.ensure(sizeof(double) * size_t(n) * R::N);
  const int total = derivTotalDoubles(s.dim);
  s.out.ensure(sizeof(double) * size_t(n) * total);
  s.outStride = n;
  s.reduceTmp.ensure(sizeof(double) * 16);
  unsigned long long* dncalc = reinterpret_cast<unsigned long long*>(s.reduceTmp.as<double>() + 8);

  SG_CHECK_HIP(hipEventRecord(s.ev[4], st));
  hipLaunchKernelGGL((k_pack<D>), dim3(nblocks(n)), dim3(kBlock), 0, st, s.perm.as<int32_t>(), n, s.pos.as<double>(),
                     s.vel.as<double>(), s.mass.as<double>(), s.rho.as<double>(), s.P.as<double>(), s.cs.as<double>(),
                     s.H.as<double>(), s.omega.as<double>(), s.hasMultipliers ? s.fCl.as<double>() : nullptr,
                     s.hasMultipliers ? s.fCq.as<double>() : nullptr, s.hasMultipliers ? s.fshear.as<double>() : nullptr,
                     s.rec.as<double>());
  SG_CHECK_HIP(hipMemsetAsync(dncalc, 0, sizeof(unsigned long long), st));
  SG_CHECK_HIP(hipEventRecord(s.ev[5], st));

  TabArgs t{s.dTable.as<double>(), s.table.n, s.table.step, s.table.extent, s.dNperh.as<double>(), s.dWsum.as<double>()};
  const size_t ldsBytes = sizeof(double) * 6 * size_t(s.table.n);
  const bool useLds = ldsBytes <= 64 * 1024;
  PassArgs p{};
  p.Cl = hp.Cl; p.Cq = hp.Cq; p.eps2 = hp.eps2; p.epsTensile = hp.epsTensile; p.nPerh = hp.nPerh;
  p.hmin = hp.hmin; p.hmax = hp.hmax;
  p.W0 = s.table.kernelValue(0.0, 1.0);
  p.WnPerh = s.table.kernelValue(1.0 / hp.nPerh, 1.0);
  p.XSPH = hp.XSPH; p.correctVelocityGradient = hp.correctVelocityGradient; p.evolveTotalEnergy = hp.evolveTotalEnergy;
  p.linearInExpansion = hp.linearInExpansion; p.quadraticInExpansion = hp.quadraticInExpansion;
  OutArgs o{};
  o.base = s.out.as<double>();
  o.n = n;
  for (int f = 0; f < SPHGPU_NUM_DERIV_FIELDS; ++f) o.off[f] = derivOffsetDoubles(s.dim, f);
  const bool tensile = hp.epsTensile != 0.0;
  const bool mult = s.hasMultipliers;
  auto launch = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nblocks(n)), dim3(kBlock), useLds ? ldsBytes : 0, st, s.rec.as<double>(),
                       s.perm.as<int32_t>(), s.offsets.as<int64_t>(), s.nbr.as<int32_t>(), n, s.nInt, t, useLds, p, o,
                       dncalc);
  };
  if (tensile && mult) launch(k_derivs_fp64<D, true, true>);
  else if (tensile) launch(k_derivs_fp64<D, true, false>);
  else if (mult) launch(k_derivs_fp64<D, false, true>);
  else launch(k_derivs_fp64<D, false, false>);
  SG_CHECK_HIP(hipGetLastError());
  SG_CHECK_HIP(hipEventRecord(s.ev[6], st));
  unsigned long long nc = 0;
  SG_CHECK_HIP(hipMemcpyAsync(&nc, dncalc, sizeof(nc), hipMemcpyDeviceToHost, st));
  SG_CHECK_HIP(hipStreamSynchronize(st));
  s.ncalc = int64_t(nc);
}

}  // namespace

void sumDensity(Sim& s) {
  SG_REQUIRE(s.neighborsValid, SPHGPU_ERR_STATE, "sum density needs a neighbour build");
  SG_REQUIRE(s.table.n > 0, SPHGPU_ERR_STATE, "no kernel table installed");
  if (s.xsStale) regatherSorted(s);
  if (s.dim == 1) sumDensityD<1>(s);
  else if (s.dim == 2) sumDensityD<2>(s);
  else sumDensityD<3>(s);
  float ms = 0.f;
  SG_CHECK_HIP(hipEventSynchronize(s.ev[3]));
  SG_CHECK_HIP(hipEventElapsedTime(&ms, s.ev[2], s.ev[3]));
  s.timing.density_ms = ms;
}

void evalDerivs(Sim& s, const sphgpu_hydro_params& p) {
  SG_REQUIRE(s.neighborsValid, SPHGPU_ERR_STATE, "evaluateDerivatives needs a neighbour build");
  SG_REQUIRE(s.table.n > 0, SPHGPU_ERR_STATE, "no kernel table installed");
  SG_REQUIRE(p.compatibleEnergyEvolution == 0, SPHGPU_ERR_UNSUPPORTED,
             "compatibleEnergyEvolution (pairAccelerations) is not produced by the device pass");
  SG_REQUIRE(p.precision == SPHGPU_PRECISION_FP64, SPHGPU_ERR_UNSUPPORTED, "precision mode not available");
  SG_REQUIRE(p.nPerh > 0.0, SPHGPU_ERR_ARG, "nPerh must be positive");
  if (s.dim == 1) evalDerivsD<1>(s, p);
  else if (s.dim == 2) evalDerivsD<2>(s, p);
  else evalDerivsD<3>(s, p);
  float ms = 0.f, msp = 0.f;
  SG_CHECK_HIP(hipEventElapsedTime(&ms, s.ev[5], s.ev[6]));
  SG_CHECK_HIP(hipEventElapsedTime(&msp, s.ev[4], s.ev[5]));
  s.timing.derivs_ms = ms;
  s.timing.pack_ms = msp;
  s.derivsValid = true;
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
