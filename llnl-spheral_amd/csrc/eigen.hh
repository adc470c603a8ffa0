// Eigenvalues of a SymTensor on the device, as SymTensor::eigenValues()
// computes them (src/Geometry/GeomSymmetricTensorInline_default.hh:2053-2110):
// 1-D the element, 2-D the quadratic formula (diagonal when |xy| < 1e-50),
// 3-D the closed form through the characteristic polynomial.  For H = h I the
// 3-D trigonometric terms are multiplied by an exact zero, so the result is
// fscale * trace/3 whatever atan2, cos and sin return.
#pragma once
#include "dgeom.hh"

namespace sphgpu {

template <int D> __device__ __forceinline__ void eigenValues(const double* h, double* e) {
  if constexpr (D == 1) {
    e[0] = h[0];
  } else if constexpr (D == 2) {
    if (dabs(h[1]) < 1.0e-50) { e[0] = h[0]; e[1] = h[2]; return; }
    const double b = h[0] + h[2];
    const double c = h[0] * h[2] - h[1] * h[1];
    const double q = 0.5 * (b + (b < 0.0 ? -1.0 : 1.0) * sqrt(dmax(0.0, b * b - 4.0 * c)));
    e[0] = q; e[1] = c / q;
  } else {
    const double onethird = 1.0 / 3.0;
    const double sqrt3 = 1.7320508075688772;   // std::sqrt(3.0), correctly rounded
    const double fscale = dmax(10.0 * 2.220446049250313e-16,
                               dmax(dabs(h[0]), dmax(dabs(h[1]), dmax(dabs(h[2]), dmax(dabs(h[3]),
                                    dmax(dabs(h[4]), dabs(h[5])))))));
    const double fscalei = 1.0 / fscale;
    const double a00 = h[0] * fscalei, a01 = h[1] * fscalei, a02 = h[2] * fscalei;
    const double a11 = h[3] * fscalei, a12 = h[4] * fscalei, a22 = h[5] * fscalei;
    const double c0 = a00 * a11 * a22 + 2.0 * a01 * a02 * a12 - a00 * a12 * a12 - a11 * a02 * a02 - a22 * a01 * a01;
    const double c1 = a00 * a11 - a01 * a01 + a00 * a22 - a02 * a02 + a11 * a22 - a12 * a12;
    const double c2 = a00 + a11 + a22;
    const double c2Div3 = c2 * onethird;
    const double aDiv3 = dmin(0.0, onethird * (c1 - c2 * c2Div3));
    const double mbDiv2 = 0.5 * (c0 + c2Div3 * (2.0 * c2Div3 * c2Div3 - c1));
    const double q = dmin(0.0, mbDiv2 * mbDiv2 + aDiv3 * aDiv3 * aDiv3);
    const double mag = sqrt(-aDiv3);
    const double angle = atan2(sqrt(-q), mbDiv2) * onethird;
    const double cs = cos(angle), sn = sin(angle);
    e[0] = fscale * (c2Div3 + 2.0 * mag * cs);
    e[1] = fscale * (c2Div3 - mag * (cs + sqrt3 * sn));
    e[2] = fscale * (c2Div3 - mag * (cs - sqrt3 * sn));
  }
}
// eigenValues().minElement() / maxElement() (std::min / std::max order)
template <int D> __device__ __forceinline__ double minEigen(const double* h) {
  double e[D];
  eigenValues<D>(h, e);
  double m = e[0];
  for (int k = 1; k < D; ++k) m = dmin(m, e[k]);
  return m;
}
template <int D> __device__ __forceinline__ double maxEigen(const double* h) {
  double e[D];
  eigenValues<D>(h, e);
  double m = e[0];
  for (int k = 1; k < D; ++k) m = dmax(m, e[k]);
  return m;
}

}  // namespace sphgpu
