// Test infrastructure: CPU oracle (see geom.hh header).
//
// Cubic B-spline kernel and the tabulated kernel built from it, restated from
//   src/Kernel/BSplineKernelInline.hh:9-100      (normalisations, W, W', W'')
//   src/Kernel/TableKernel.cc:211-305             (construction, step = extent/(n-1))
//   src/Kernel/TableKernel.cc:469-506             (setParabolicCoeffs: C = inv(A).B with
//                                                  A = [[0,0,1],[1,1,1],[4,2,1]]; the inverse is
//                                                  exact, rows [.5,-1,.5], [-1.5,2,-.5], [1,-0,-0])
//   src/Kernel/TableKernel.cc:511-548, 23-93      (nperh <-> Wsum tables)
//   src/Kernel/TableKernel.cc:399-429             (equivalentNodesPerSmoothingScale)
//   src/Kernel/TableKernelInline.hh:11-78, 262-294 (lookups, lowerBound, parabolicInterp)
//   src/Utilities/bisectSearch.hh:20-50
#pragma once
#include <cmath>
#include <vector>
#include <algorithm>
#include "geom.hh"

namespace orc {

template <int D> inline double bsplineNorm();
template <> inline double bsplineNorm<1>() { return 2.0 / 3.0; }
template <> inline double bsplineNorm<2>() { return 10.0 / (7.0 * M_PI); }
template <> inline double bsplineNorm<3>() { return 1.0 / M_PI; }

template <int D> inline double bsplineW(double eta, double Hdet) {
  const double A = bsplineNorm<D>();
  if (eta < 1.0) {
    const double eta2 = eta * eta;
    return A * Hdet * (1.0 - 1.5 * eta2 + 0.75 * eta2 * eta);
  } else if (eta < 2.0) {
    const double t = 2.0 - eta;
    return A * Hdet * 0.25 * (t * t * t);
  }
  return 0.0;
}
template <int D> inline double bsplineGradW(double eta, double Hdet) {
  const double A = bsplineNorm<D>();
  if (eta < 1.0) return -A * Hdet * (3.0 - 2.25 * eta) * eta;
  if (eta < 2.0) { const double t = 2.0 - eta; return -A * Hdet * 0.75 * (t * t); }
  return 0.0;
}
template <int D> inline double bsplineGrad2W(double eta, double Hdet) {
  const double A = bsplineNorm<D>();
  if (eta < 1.0) return -A * Hdet * (3.0 - 4.5 * eta);
  if (eta < 2.0) return A * Hdet * 1.5 * (2.0 - eta);
  return 0.0;
}

struct Table {
  int dim = 3;
  int n = 0;
  double extent = 2.0;
  double step = 0.0;
  std::vector<double> Ak, Bk, Ck, Ag, Bg, Cg, Ag2, Bg2, Cg2;
  std::vector<double> nperh, wsum;
  double minNperh = 0.25, maxNperh = 10.0;

  int lowerBound(double eta) const { return std::min(n - 1, int(eta / step)); }
  double interp(double eta, const std::vector<double>& a, const std::vector<double>& b, const std::vector<double>& c) const {
    const int i0 = std::min(n - 3, lowerBound(eta));
    const int i1 = i0 + 1;
    const double x = eta / step - i0;
    return a[i1] + b[i1] * x + c[i1] * x * x;
  }
  double kernelValue(double eta, double Hdet) const { return eta < extent ? Hdet * interp(eta, Ak, Bk, Ck) : 0.0; }
  double gradValue(double eta, double Hdet) const { return eta < extent ? Hdet * interp(eta, Ag, Bg, Cg) : 0.0; }
  double grad2Value(double eta, double Hdet) const { return eta < extent ? Hdet * interp(eta, Ag2, Bg2, Cg2) : 0.0; }
  void kernelAndGradValue(double eta, double Hdet, double& W, double& gW) const {
    if (eta < extent) {
      const int i0 = std::min(n - 3, lowerBound(eta));
      const int i1 = i0 + 1;
      const double x = eta / step - i0;
      W = Hdet * (Ak[i1] + Bk[i1] * x + Ck[i1] * x * x);
      gW = Hdet * (Ag[i1] + Bg[i1] * x + Cg[i1] * x * x);
    } else {
      W = 0.0; gW = 0.0;
    }
  }

  // bisectSearch over an ascending table.
  static int bisect(const std::vector<double>& t, double val) {
    const int nn = int(t.size());
    if (nn > 1) {
      const bool ascnd = (t[nn - 1] >= t[0]);
      int jl = -1, ju = nn;
      while (ju - jl > 1) {
        const int jm = (ju + jl) / 2;
        if ((val >= t[jm]) == ascnd) jl = jm; else ju = jm;
      }
      return jl;
    } else if (nn == 1) {
      return val <= t[0] ? -1 : 0;
    }
    return -1;
  }
  double equivalentNodesPerSmoothingScale(double Wsum) const {
    const int lb = bisect(wsum, Wsum);
    const int ub = lb + 1;
    if (lb == -1) return nperh[0];
    if (ub == n) return nperh[n - 1];
    return std::min(nperh[ub], std::max(nperh[lb], nperh[lb] + (Wsum - wsum[lb]) / (wsum[ub] - wsum[lb]) * (nperh[ub] - nperh[lb])));
  }
};

inline void setParabolicCoeffs(int n, const std::vector<double>& t, std::vector<double>& a, std::vector<double>& b, std::vector<double>& c) {
  a.assign(n, 0.0); b.assign(n, 0.0); c.assign(n, 0.0);
  for (int i0 = 0; i0 < n - 2; ++i0) {
    const int i1 = i0 + 1, i2 = i0 + 2;
    const double t0 = t[i0], t1 = t[i1], t2 = t[i2];
    c[i1] = (0.5 * t0 + (-1.0) * t1) + 0.5 * t2;
    b[i1] = ((-1.5) * t0 + 2.0 * t1) + (-0.5) * t2;
    a[i1] = (1.0 * t0 + (-0.0) * t1) + (-0.0) * t2;
  }
}

template <int D> double sumKernelValues(const Table& W, double deta);
template <> inline double sumKernelValues<1>(const Table& W, double deta) {
  double result = 0.0, etax = deta;
  while (etax < W.extent) { result += 2.0 * std::fabs(W.gradValue(etax, 1.0)); etax += deta; }
  return result;
}
template <> inline double sumKernelValues<2>(const Table& W, double deta) {
  double result = 0.0, etay = 0.0;
  while (etay < W.extent) {
    double etax = 0.0;
    while (etax < W.extent) {
      const Vec<2> eta{{etax, etay}};
      double d = std::fabs(W.gradValue(mag(eta), 1.0));
      if (distinctlyGreaterThan(etax, 0.0)) d *= 2.0;
      if (distinctlyGreaterThan(etay, 0.0)) d *= 2.0;
      if (fuzzyEqual(mag(eta), 0.0)) d *= 0.0;
      result += d;
      etax += deta;
    }
    etay += deta;
  }
  return std::sqrt(result);
}
template <> inline double sumKernelValues<3>(const Table& W, double deta) {
  double result = 0.0, etaz = 0.0;
  while (etaz < W.extent) {
    double etay = 0.0;
    while (etay < W.extent) {
      double etax = 0.0;
      while (etax < W.extent) {
        const Vec<3> eta{{etax, etay, etaz}};
        double d = std::fabs(W.gradValue(mag(eta), 1.0));
        if (distinctlyGreaterThan(etax, 0.0)) d *= 2.0;
        if (distinctlyGreaterThan(etay, 0.0)) d *= 2.0;
        if (distinctlyGreaterThan(etaz, 0.0)) d *= 2.0;
        if (fuzzyEqual(mag(eta), 0.0)) d *= 0.0;
        result += d;
        etax += deta;
      }
      etay += deta;
    }
    etaz += deta;
  }
  return rootnu<3>(result);
}

template <int D> Table buildBSplineTable(int numPoints) {
  Table T;
  T.dim = D;
  T.n = numPoints;
  const double hmult = 1.0;
  T.extent = hmult * 2.0;
  T.step = T.extent / (numPoints - 1);
  const double correction = 1.0 / pownu<D>(hmult);
  const double deta = T.step / hmult;
  std::vector<double> kv(numPoints), gv(numPoints), g2v(numPoints);
  for (int i = 0; i < numPoints; ++i) {
    kv[i] = correction * bsplineW<D>(i * deta, 1.0);
    gv[i] = correction * bsplineGradW<D>(i * deta, 1.0);
    g2v[i] = correction * bsplineGrad2W<D>(i * deta, 1.0);
  }
  setParabolicCoeffs(numPoints, kv, T.Ak, T.Bk, T.Ck);
  setParabolicCoeffs(numPoints, gv, T.Ag, T.Bg, T.Cg);
  setParabolicCoeffs(numPoints, g2v, T.Ag2, T.Bg2, T.Cg2);
  T.wsum.assign(numPoints, 0.0);
  T.nperh.assign(numPoints, 0.0);
  const double dnperh = (T.maxNperh - T.minNperh) / (numPoints - 1);
  for (int i = 0; i != numPoints; ++i) {
    const double nperh = T.minNperh + i * dnperh;
    T.nperh[i] = nperh;
    T.wsum[i] = sumKernelValues<D>(T, 1.0 / nperh);
  }
  return T;
}

}  // namespace orc
