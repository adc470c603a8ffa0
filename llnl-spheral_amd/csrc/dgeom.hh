// Device-side Dim<nu> geometry for the SPH derivative pass.
//
// Every FP64 expression keeps the reference's association order so that the
// FP64 path reproduces the reference sums bit for bit (the library is compiled
// with -ffp-contract=off, so no multiply-add is fused):
//   Vector ops      src/Geometry/GeomVectorInline_default.hh:398-574, 804-969
//   SymTensor ops   src/Geometry/GeomSymmetricTensorInline_default.hh:898-945,
//                   1596-1661 (Trace, Determinant, dot(Vector))
//   Tensor ops      src/Geometry/GeomTensorInline_default.hh:885-937, 1286-1340,
//                   1800-1822 (Inverse), 1861-1895, 1938-1957
// SymTensor storage is xx,xy,xz,yy,yz,zz (GeomSymmetricTensorBase.hh:64-66);
// Tensor storage is row major.
#pragma once
#include <hip/hip_runtime.h>

namespace sphgpu {

#define SG_HD __host__ __device__ __forceinline__

template <int D> struct NSym { static constexpr int value = D * (D + 1) / 2; };

template <int D> struct V { double c[D]; };
template <int D> struct T { double c[D * D]; };
template <int D> struct S { double c[NSym<D>::value]; };

SG_HD double dmax(double a, double b) { return (a < b) ? b : a; }   // std::max
SG_HD double dmin(double a, double b) { return (b < a) ? b : a; }   // std::min
SG_HD double dabs(double a) { return a < 0.0 ? -a : (a == 0.0 ? 0.0 : a); }

template <int D> SG_HD V<D> vzero() { V<D> r; for (int k = 0; k < D; ++k) r.c[k] = 0.0; return r; }
template <int D> SG_HD T<D> tzero() { T<D> r; for (int k = 0; k < D * D; ++k) r.c[k] = 0.0; return r; }
template <int D> SG_HD S<D> szero() { S<D> r; for (int k = 0; k < NSym<D>::value; ++k) r.c[k] = 0.0; return r; }

template <int D> SG_HD V<D> vsub(const V<D>& a, const V<D>& b) { V<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] - b.c[k]; return r; }
template <int D> SG_HD V<D> vscale(const V<D>& a, double s) { V<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] * s; return r; }
template <int D> SG_HD double vdot(const V<D>& a, const V<D>& b) {
  double r = a.c[0] * b.c[0];
  for (int k = 1; k < D; ++k) r = r + a.c[k] * b.c[k];
  return r;
}
template <int D> SG_HD double vmag2(const V<D>& a) { return vdot(a, a); }
template <int D> SG_HD double vmag(const V<D>& a) {
  if constexpr (D == 1) return dabs(a.c[0]);
  else return __builtin_sqrt(vmag2(a));
}
// Vector/scalar: 1-D divides, higher dimensions multiply by 1/s.
template <int D> SG_HD V<D> vdivs(const V<D>& a, double s) {
  V<D> r;
  if constexpr (D == 1) { r.c[0] = a.c[0] / s; }
  else { const double inv = 1.0 / s; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] * inv; }
  return r;
}
template <int D> SG_HD V<D> vunit(const V<D>& a, double magnitude) {
  if (magnitude > 1.0e-50) return vdivs(a, magnitude);
  V<D> r = vzero<D>(); r.c[0] = 1.0; return r;
}

// SymTensor.Vector, row by row, left to right.
template <int D> SG_HD V<D> sdot(const S<D>& H, const V<D>& r) {
  V<D> o;
  if constexpr (D == 1) {
    o.c[0] = H.c[0] * r.c[0];
  } else if constexpr (D == 2) {
    o.c[0] = H.c[0] * r.c[0] + H.c[1] * r.c[1];
    o.c[1] = H.c[1] * r.c[0] + H.c[2] * r.c[1];
  } else {
    o.c[0] = H.c[0] * r.c[0] + H.c[1] * r.c[1] + H.c[2] * r.c[2];
    o.c[1] = H.c[1] * r.c[0] + H.c[3] * r.c[1] + H.c[4] * r.c[2];
    o.c[2] = H.c[2] * r.c[0] + H.c[4] * r.c[1] + H.c[5] * r.c[2];
  }
  return o;
}
template <int D> SG_HD double sdet(const S<D>& H) {
  if constexpr (D == 1) return H.c[0];
  else if constexpr (D == 2) return H.c[0] * H.c[2] - H.c[1] * H.c[1];
  else {
    const double xx = H.c[0], xy = H.c[1], xz = H.c[2], yy = H.c[3], yz = H.c[4], zz = H.c[5];
    return xx * yy * zz + xy * yz * xz + xz * xy * yz - xx * yz * yz - xy * xy * zz - xz * yy * xz;
  }
}
// Storage index of diagonal element i.
template <int D> SG_HD constexpr int sdiag(int i) { return i * D - (i * (i - 1)) / 2; }

template <int D> SG_HD double ttrace(const T<D>& a) { double s = a.c[0]; for (int k = 1; k < D; ++k) s = s + a.c[k * D + k]; return s; }
template <int D> SG_HD double tdet(const T<D>& a) {
  if constexpr (D == 1) return a.c[0];
  else if constexpr (D == 2) return a.c[0] * a.c[3] - a.c[1] * a.c[2];
  else {
    const double xx = a.c[0], xy = a.c[1], xz = a.c[2], yx = a.c[3], yy = a.c[4], yz = a.c[5], zx = a.c[6], zy = a.c[7], zz = a.c[8];
    return xx * yy * zz + xy * yz * zx + xz * yx * zy - xx * yz * zy - xy * yx * zz - xz * yy * zx;
  }
}
template <int D> SG_HD T<D> tinv(const T<D>& a) {
  T<D> r;
  if constexpr (D == 1) {
    r.c[0] = 1.0 / a.c[0];
  } else if constexpr (D == 2) {
    const double inv = 1.0 / tdet(a);
    r.c[0] = a.c[3] * inv; r.c[1] = -a.c[1] * inv; r.c[2] = -a.c[2] * inv; r.c[3] = a.c[0] * inv;
  } else {
    const double xx = a.c[0], xy = a.c[1], xz = a.c[2], yx = a.c[3], yy = a.c[4], yz = a.c[5], zx = a.c[6], zy = a.c[7], zz = a.c[8];
    const double inv = 1.0 / tdet(a);
    r.c[0] = (yy * zz - yz * zy) * inv; r.c[1] = (xz * zy - xy * zz) * inv; r.c[2] = (xy * yz - xz * yy) * inv;
    r.c[3] = (yz * zx - yx * zz) * inv; r.c[4] = (xx * zz - xz * zx) * inv; r.c[5] = (xz * yx - xx * yz) * inv;
    r.c[6] = (yx * zy - yy * zx) * inv; r.c[7] = (xy * zx - xx * zy) * inv; r.c[8] = (xx * yy - xy * yx) * inv;
  }
  return r;
}
template <int D> SG_HD T<D> tdot(const T<D>& a, const T<D>& b) {
  T<D> r;
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      double s = a.c[i * D] * b.c[j];
      for (int k = 1; k < D; ++k) s = s + a.c[i * D + k] * b.c[k * D + j];
      r.c[i * D + j] = s;
    }
  return r;
}
// Tensor /= s: 1-D divides, higher dimensions multiply by 1/s.
template <int D> SG_HD void tdivEq(T<D>& a, double s) {
  if constexpr (D == 1) { a.c[0] = a.c[0] / s; }
  else { const double inv = 1.0 / s; for (int k = 0; k < D * D; ++k) a.c[k] = a.c[k] * inv; }
}
template <int D> SG_HD S<D> sdivs(const S<D>& a, double s) {
  S<D> r;
  if constexpr (D == 1) { r.c[0] = a.c[0] / s; }
  else { const double inv = 1.0 / s; for (int k = 0; k < NSym<D>::value; ++k) r.c[k] = a.c[k] * inv; }
  return r;
}

// nu-th root / power (src/Geometry/Dimension.hh:53-98; 3-D root is
// sgn(x)*pow(|x|, 0.3333333333333333), src/Utilities/FastMath.hh:148-156).
template <int D> __device__ __forceinline__ double rootnu(double x) {
  if constexpr (D == 1) return x;
  else if constexpr (D == 2) return __builtin_sqrt(x);
  else return (x < 0.0 ? -1.0 : 1.0) * pow(dabs(x), 0.3333333333333333);
}
template <int D> SG_HD constexpr double pownu2() { return D == 1 ? 2.0 : (D == 2 ? 4.0 : 8.0); }

// XCD-aware block order.  The dispatcher deals workgroups round-robin to the 8
// XCDs of an MI355X, each with its own L2; remapping so that XCD x runs the
// logical blocks [x*per, (x+1)*per) keeps spatially adjacent blocks (and their
// shared neighbours) in one L2.  Launch kXcds*per blocks; logical ids >= nb
// have no work.
constexpr int kXcds = 8;
__host__ __forceinline__ int64_t xcdGrid(int64_t nb) { return kXcds * ((nb + kXcds - 1) / kXcds); }
__device__ __forceinline__ int64_t xcdBlock(int64_t nb) {
  const int64_t per = (nb + kXcds - 1) / kXcds;
  return int64_t(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

// src/Utilities/SpheralFunctions.hh:20-23; src/Utilities/safeInv.hh:16-27.
SG_HD bool fuzzyEqual(double l, double r, double fuzz = 1.0e-15) {
  return dabs(l - r) / dmax(1.0, dabs(l) + dabs(r)) < fuzz;
}
SG_HD double safeInv(double x, double fuzz = 1.0e-30) { return x / (x * x + fuzz); }
SG_HD double safeInvVar(double x, double fuzz = 1.0e-30) { return (x < 0.0 ? -1.0 : 1.0) / dmax(fuzz, dabs(x)); }

}  // namespace sphgpu
