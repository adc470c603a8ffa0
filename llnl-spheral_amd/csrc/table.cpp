// Host construction of TableKernel(BSplineKernel<nu>(), n) for the device pass.
//
//   BSplineKernel<nu>::kernelValue / gradValue   src/Kernel/BSplineKernelInline.hh:9-83
//   TableKernel ctor (step = extent/(n-1))       src/Kernel/TableKernel.cc:211-305
//   setParabolicCoeffs                           src/Kernel/TableKernel.cc:469-506
//     The fit matrix A = [[0,0,1],[1,1,1],[4,2,1]] has the exact inverse
//     rows [1/2,-1,1/2], [-3/2,2,-1/2], [1,0,0]; the coefficients are those rows
//     dotted with (t0,t1,t2), summed left to right.
//   setNperhValues / sumKernelValues             src/Kernel/TableKernel.cc:23-93, 511-548
//   parabolicInterp / lowerBound                 src/Kernel/TableKernelInline.hh:262-294
#include <cmath>
#include <algorithm>
#include "internal.hh"

namespace sphgpu {
namespace {

double bnorm(int dim) { return dim == 1 ? 2.0 / 3.0 : (dim == 2 ? 10.0 / (7.0 * M_PI) : 1.0 / M_PI); }

double bW(int dim, double eta) {
  const double A = bnorm(dim), Hdet = 1.0;
  if (eta < 1.0) { const double e2 = eta * eta; return A * Hdet * (1.0 - 1.5 * e2 + 0.75 * e2 * eta); }
  if (eta < 2.0) { const double t = 2.0 - eta; return A * Hdet * 0.25 * (t * t * t); }
  return 0.0;
}
double bdW(int dim, double eta) {
  const double A = bnorm(dim), Hdet = 1.0;
  if (eta < 1.0) return -A * Hdet * (3.0 - 2.25 * eta) * eta;
  if (eta < 2.0) { const double t = 2.0 - eta; return -A * Hdet * 0.75 * (t * t); }
  return 0.0;
}

void fit(int n, const std::vector<double>& t, std::vector<double>& a, std::vector<double>& b, std::vector<double>& c) {
  a.assign(n, 0.0); b.assign(n, 0.0); c.assign(n, 0.0);
  for (int i = 1; i + 1 < n; ++i) {
    const double t0 = t[i - 1], t1 = t[i], t2 = t[i + 1];
    c[i] = (0.5 * t0 + (-1.0) * t1) + 0.5 * t2;
    b[i] = ((-1.5) * t0 + 2.0 * t1) + (-0.5) * t2;
    a[i] = (1.0 * t0 + (-0.0) * t1) + (-0.0) * t2;
  }
}

bool fuzzyZero(double x) { return std::fabs(x) / std::max(1.0, std::fabs(x)) < 1.0e-15; }
bool distinctlyPositive(double x) { return x > 0.0 && !fuzzyZero(x); }

// Sum of |dW/deta| over the lattice of spacing deta (TableKernel.cc:23-93).
double latticeGradSum(const TableHost& T, int dim, double deta) {
  double result = 0.0;
  if (dim == 1) {
    for (double ex = deta; ex < T.extent; ex += deta) result += 2.0 * std::fabs(T.gradValue(ex, 1.0));
    return result;
  }
  if (dim == 2) {
    for (double ey = 0.0; ey < T.extent; ey += deta)
      for (double ex = 0.0; ex < T.extent; ex += deta) {
        const double r = std::sqrt(ex * ex + ey * ey);
        double d = std::fabs(T.gradValue(r, 1.0));
        if (distinctlyPositive(ex)) d *= 2.0;
        if (distinctlyPositive(ey)) d *= 2.0;
        if (fuzzyZero(r)) d *= 0.0;
        result += d;
      }
    return std::sqrt(result);
  }
  for (double ez = 0.0; ez < T.extent; ez += deta)
    for (double ey = 0.0; ey < T.extent; ey += deta)
      for (double ex = 0.0; ex < T.extent; ex += deta) {
        const double r = std::sqrt(ex * ex + ey * ey + ez * ez);
        double d = std::fabs(T.gradValue(r, 1.0));
        if (distinctlyPositive(ex)) d *= 2.0;
        if (distinctlyPositive(ey)) d *= 2.0;
        if (distinctlyPositive(ez)) d *= 2.0;
        if (fuzzyZero(r)) d *= 0.0;
        result += d;
      }
  return (result < 0.0 ? -1.0 : 1.0) * std::pow(std::fabs(result), 0.3333333333333333);
}

}  // namespace

double TableHost::kernelValue(double eta, double Hdet) const {
  if (!(eta < extent)) return 0.0;
  const double q = eta / step;
  const int i0 = std::min(n - 3, std::min(n - 1, int(q)));
  const int i1 = i0 + 1;
  const double x = q - i0;
  return Hdet * (Ak[i1] + Bk[i1] * x + Ck[i1] * x * x);
}
double TableHost::gradValue(double eta, double Hdet) const {
  if (!(eta < extent)) return 0.0;
  const double q = eta / step;
  const int i0 = std::min(n - 3, std::min(n - 1, int(q)));
  const int i1 = i0 + 1;
  const double x = q - i0;
  return Hdet * (Ag[i1] + Bg[i1] * x + Cg[i1] * x * x);
}

void buildBSplineTable(int dim, int npoints, TableHost& T) {
  T.n = npoints;
  T.extent = 1.0 * 2.0;
  T.step = T.extent / (npoints - 1);
  std::vector<double> kv(npoints), gv(npoints);
  for (int i = 0; i < npoints; ++i) {
    kv[i] = 1.0 * bW(dim, i * T.step);
    gv[i] = 1.0 * bdW(dim, i * T.step);
  }
  fit(npoints, kv, T.Ak, T.Bk, T.Ck);
  fit(npoints, gv, T.Ag, T.Bg, T.Cg);
  const double nmin = 0.25, nmax = 10.0;
  const double dn = (nmax - nmin) / (npoints - 1);
  T.nperh.assign(npoints, 0.0);
  T.wsum.assign(npoints, 0.0);
  for (int i = 0; i < npoints; ++i) {
    T.nperh[i] = nmin + i * dn;
    T.wsum[i] = latticeGradSum(T, dim, 1.0 / T.nperh[i]);
  }
}

}  // namespace sphgpu
