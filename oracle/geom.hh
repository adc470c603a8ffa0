// Test infrastructure: CPU oracle for the SPH derivative pass.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it;
// it is never part of the product path.
//
// Geometry primitives restated from the reference so that every floating-point
// operation happens in the same order (bit-exact neighbour sets and sums depend
// on it).  Compiled with -ffp-contract=off: no FMA contraction anywhere.
//
//   Vector  : src/Geometry/GeomVectorInline_default.hh:374-574 (ops), 804-820 (dot),
//             855-873 (dyad), 879-910 (selfdyad), 918-921 (unitVector),
//             926-969 (magnitude, magnitude2)
//   SymTensor (storage xx,xy,xz,yy,yz,zz as GeomSymmetricTensorBase.hh:64-66):
//             src/Geometry/GeomSymmetricTensorInline_default.hh:898-945 (scalar ops),
//             1529-1557 (Inverse), 1596-1605 (Trace), 1613-1634 (Determinant),
//             1642-1661 (dot with Vector)
//   Tensor  (row major xx,xy,xz,yx,...): src/Geometry/GeomTensorInline_default.hh:
//             885-937 (scalar ops), 965-1049 (+=,-=), 1286-1340 (*=,/=),
//             1800-1822 (Inverse), 1861-1870 (Trace), 1878-1895 (Determinant),
//             1938-1957 (dot with Tensor)
#pragma once
#include <cmath>

namespace orc {

template <int D> struct Vec { double c[D]; };
template <int D> struct Sym;   // D*(D+1)/2 components
template <int D> struct Ten { double c[D * D]; };
template <> struct Sym<1> { double xx; };
template <> struct Sym<2> { double xx, xy, yy; };
template <> struct Sym<3> { double xx, xy, xz, yy, yz, zz; };

template <int D> constexpr int nsym() { return D * (D + 1) / 2; }

template <int D> inline Vec<D> vzero() { Vec<D> r; for (int k = 0; k < D; ++k) r.c[k] = 0.0; return r; }
template <int D> inline Ten<D> tzero() { Ten<D> r; for (int k = 0; k < D * D; ++k) r.c[k] = 0.0; return r; }
template <int D> inline Sym<D> szero();
template <> inline Sym<1> szero<1>() { return Sym<1>{0.0}; }
template <> inline Sym<2> szero<2>() { return Sym<2>{0.0, 0.0, 0.0}; }
template <> inline Sym<3> szero<3>() { return Sym<3>{0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; }

template <int D> inline Vec<D> sub(const Vec<D>& a, const Vec<D>& b) { Vec<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] - b.c[k]; return r; }
template <int D> inline Vec<D> add(const Vec<D>& a, const Vec<D>& b) { Vec<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] + b.c[k]; return r; }
// Vector*scalar: each component times the scalar.
template <int D> inline Vec<D> scale(const Vec<D>& a, double s) { Vec<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] * s; return r; }
template <int D> inline void addTo(Vec<D>& a, const Vec<D>& b) { for (int k = 0; k < D; ++k) a.c[k] += b.c[k]; }
template <int D> inline void subFrom(Vec<D>& a, const Vec<D>& b) { for (int k = 0; k < D; ++k) a.c[k] -= b.c[k]; }

// dot / magnitude2: left-to-right sums.
template <int D> inline double dot(const Vec<D>& a, const Vec<D>& b) {
  double r = a.c[0] * b.c[0];
  for (int k = 1; k < D; ++k) r = r + a.c[k] * b.c[k];
  return r;
}
template <int D> inline double mag2(const Vec<D>& a) { return dot(a, a); }
template <int D> inline double mag(const Vec<D>& a) { return std::sqrt(mag2(a)); }
template <> inline double mag<1>(const Vec<1>& a) { return std::fabs(a.c[0]); }

// Vector/scalar: 1-D divides, 2-D/3-D multiply by 1/val (GeomVectorInline_default.hh:547-574).
template <int D> inline Vec<D> divs(const Vec<D>& a, double s) {
  const double inv = 1.0 / s;
  Vec<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] * inv; return r;
}
template <> inline Vec<1> divs<1>(const Vec<1>& a, double s) { return Vec<1>{{a.c[0] / s}}; }

template <int D> inline Vec<D> unitVector(const Vec<D>& a) {
  const double m = mag(a);
  if (m > 1.0e-50) return divs(a, m);
  Vec<D> r = vzero<D>(); r.c[0] = 1.0; return r;
}

template <int D> inline Ten<D> dyad(const Vec<D>& a, const Vec<D>& b) {
  Ten<D> r; for (int i = 0; i < D; ++i) for (int j = 0; j < D; ++j) r.c[i * D + j] = a.c[i] * b.c[j]; return r;
}
template <int D> inline Ten<D> tscale(const Ten<D>& a, double s) { Ten<D> r; for (int k = 0; k < D * D; ++k) r.c[k] = a.c[k] * s; return r; }
template <int D> inline void taddTo(Ten<D>& a, const Ten<D>& b) { for (int k = 0; k < D * D; ++k) a.c[k] += b.c[k]; }
template <int D> inline void tsubFrom(Ten<D>& a, const Ten<D>& b) { for (int k = 0; k < D * D; ++k) a.c[k] -= b.c[k]; }
template <int D> inline Ten<D> tneg(const Ten<D>& a) { Ten<D> r; for (int k = 0; k < D * D; ++k) r.c[k] = -a.c[k]; return r; }
// Tensor/=scalar (1-D divides, else multiply by inverse).
template <int D> inline void tdivEq(Ten<D>& a, double s) { const double inv = 1.0 / s; for (int k = 0; k < D * D; ++k) a.c[k] *= inv; }
template <> inline void tdivEq<1>(Ten<1>& a, double s) { a.c[0] /= s; }

// Tensor.Tensor, Trace, Determinant, Inverse.
template <int D> inline Ten<D> tdot(const Ten<D>& a, const Ten<D>& b) {
  Ten<D> r;
  for (int i = 0; i < D; ++i) for (int j = 0; j < D; ++j) {
    double s = a.c[i * D + 0] * b.c[0 * D + j];
    for (int k = 1; k < D; ++k) s = s + a.c[i * D + k] * b.c[k * D + j];
    r.c[i * D + j] = s;
  }
  return r;
}
template <int D> inline double ttrace(const Ten<D>& a) { double s = a.c[0]; for (int k = 1; k < D; ++k) s = s + a.c[k * D + k]; return s; }
inline double tdet(const Ten<1>& a) { return a.c[0]; }
inline double tdet(const Ten<2>& a) { return a.c[0] * a.c[3] - a.c[1] * a.c[2]; }
inline double tdet(const Ten<3>& a) {
  const double xx = a.c[0], xy = a.c[1], xz = a.c[2], yx = a.c[3], yy = a.c[4], yz = a.c[5], zx = a.c[6], zy = a.c[7], zz = a.c[8];
  return ((xx * yy * zz + xy * yz * zx + xz * yx * zy - xx * yz * zy - xy * yx * zz - xz * yy * zx));
}
inline Ten<1> tinv(const Ten<1>& a) { return Ten<1>{{1.0 / a.c[0]}}; }
inline Ten<2> tinv(const Ten<2>& a) {
  Ten<2> r{{a.c[3], -a.c[1], -a.c[2], a.c[0]}};
  const double d = tdet(a);            // Tensor/double: multiply by 1/d (GeomTensorInline_default.hh:921-926)
  const double inv = 1.0 / d;
  for (int k = 0; k < 4; ++k) r.c[k] = r.c[k] * inv;
  return r;
}
inline Ten<3> tinv(const Ten<3>& a) {
  const double xx = a.c[0], xy = a.c[1], xz = a.c[2], yx = a.c[3], yy = a.c[4], yz = a.c[5], zx = a.c[6], zy = a.c[7], zz = a.c[8];
  Ten<3> r{{yy * zz - yz * zy, xz * zy - xy * zz, xy * yz - xz * yy,
            yz * zx - yx * zz, xx * zz - xz * zx, xz * yx - xx * yz,
            yx * zy - yy * zx, xy * zx - xx * zy, xx * yy - xy * yx}};
  const double inv = 1.0 / tdet(a);
  for (int k = 0; k < 9; ++k) r.c[k] = r.c[k] * inv;
  return r;
}

// SymTensor.Vector (GeomSymmetricTensorInline_default.hh:1642-1661).
inline Vec<1> sdot(const Sym<1>& H, const Vec<1>& r) { return Vec<1>{{H.xx * r.c[0]}}; }
inline Vec<2> sdot(const Sym<2>& H, const Vec<2>& r) {
  return Vec<2>{{H.xx * r.c[0] + H.xy * r.c[1], H.xy * r.c[0] + H.yy * r.c[1]}};
}
inline Vec<3> sdot(const Sym<3>& H, const Vec<3>& r) {
  return Vec<3>{{H.xx * r.c[0] + H.xy * r.c[1] + H.xz * r.c[2],
                 H.xy * r.c[0] + H.yy * r.c[1] + H.yz * r.c[2],
                 H.xz * r.c[0] + H.yz * r.c[1] + H.zz * r.c[2]}};
}
inline double sdet(const Sym<1>& H) { return H.xx; }
inline double sdet(const Sym<2>& H) { return H.xx * H.yy - H.xy * H.xy; }
inline double sdet(const Sym<3>& H) {
  return ((H.xx * H.yy * H.zz + H.xy * H.yz * H.xz + H.xz * H.xy * H.yz -
           H.xx * H.yz * H.yz - H.xy * H.xy * H.zz - H.xz * H.yy * H.xz));
}
inline double strace(const Sym<1>& H) { return H.xx; }
inline double strace(const Sym<2>& H) { return H.xx + H.yy; }
inline double strace(const Sym<3>& H) { return H.xx + H.yy + H.zz; }

// Generic component access for SymTensor (storage order).
template <int D> inline double* sptr(Sym<D>& s) { return &s.xx; }
template <int D> inline const double* sptr(const Sym<D>& s) { return &s.xx; }
template <int D> inline Sym<D> sscale(const Sym<D>& a, double s) { Sym<D> r = a; double* p = sptr(r); for (int k = 0; k < nsym<D>(); ++k) p[k] *= s; return r; }
template <int D> inline void saddTo(Sym<D>& a, const Sym<D>& b) { double* p = sptr(a); const double* q = sptr(b); for (int k = 0; k < nsym<D>(); ++k) p[k] += q[k]; }
template <int D> inline Sym<D> sneg(const Sym<D>& a) { Sym<D> r = a; double* p = sptr(r); for (int k = 0; k < nsym<D>(); ++k) p[k] = -p[k]; return r; }
// SymTensor/=scalar and SymTensor/scalar (1-D divides, else multiply by inverse).
template <int D> inline Sym<D> sdivs(const Sym<D>& a, double s) { const double inv = 1.0 / s; return sscale(a, inv); }
template <> inline Sym<1> sdivs<1>(const Sym<1>& a, double s) { return Sym<1>{a.xx / s}; }

// Vector.selfdyad (GeomVectorInline_default.hh:879-910).
inline Sym<1> selfdyad(const Vec<1>& a) { return Sym<1>{a.c[0] * a.c[0]}; }
inline Sym<2> selfdyad(const Vec<2>& a) { return Sym<2>{a.c[0] * a.c[0], a.c[0] * a.c[1], a.c[1] * a.c[1]}; }
inline Sym<3> selfdyad(const Vec<3>& a) {
  return Sym<3>{a.c[0] * a.c[0], a.c[0] * a.c[1], a.c[0] * a.c[2], a.c[1] * a.c[1], a.c[1] * a.c[2], a.c[2] * a.c[2]};
}
// Identity * scalar: diagonal elements 1*s, off-diagonals 0*s.
template <int D> inline Sym<D> sidentity_times(double s) {
  Sym<D> r = szero<D>(); double* p = sptr(r);
  int k = 0;
  for (int i = 0; i < D; ++i) for (int j = i; j < D; ++j, ++k) p[k] = (i == j ? 1.0 : 0.0) * s;
  return r;
}
// Diagonal component index in storage order.
template <int D> inline int sdiag(int i) { return i * D - (i * (i - 1)) / 2; }

// Dimension<nu> helpers (src/Geometry/Dimension.hh:53-98).
template <int D> inline double rootnu(double x);
template <> inline double rootnu<1>(double x) { return x; }
template <> inline double rootnu<2>(double x) { return std::sqrt(x); }
// CubeRootHalley2 non-Intel branch: sgn(d)*pow(|d|, 0.3333333333333333) (src/Utilities/FastMath.hh:148-156).
template <> inline double rootnu<3>(double x) { return (x < 0.0 ? -1.0 : 1.0) * std::pow(std::fabs(x), 0.3333333333333333); }
template <int D> inline double pownu(double x);
template <> inline double pownu<1>(double x) { return x; }
template <> inline double pownu<2>(double x) { return x * x; }
template <> inline double pownu<3>(double x) { return x * x * x; }

// src/Utilities/SpheralFunctions.hh:20-23, 44-47, 57-59; src/Utilities/safeInv.hh:16-27.
inline bool fuzzyEqual(double l, double r, double fuzz = 1.0e-15) {
  return std::fabs(l - r) / std::fmax(1.0, std::fabs(l) + std::fabs(r)) < fuzz;
}
inline bool distinctlyGreaterThan(double l, double r, double fuzz = 1.0e-15) { return l > r && !fuzzyEqual(l, r, fuzz); }
inline double sgn(double x) { return x < 0.0 ? -1.0 : 1.0; }
inline double safeInv(double x, double fuzz = 1.0e-30) { return x / (x * x + fuzz); }
inline double safeInvVar(double x, double fuzz = 1.0e-30) { return sgn(x) / std::fmax(fuzz, std::fabs(x)); }

}  // namespace orc
