// dual.hpp -- forward-mode derivatives for the FP64 analytic form factors and structure factor.
//
// dd is a real number with one directional derivative (v + e d, e^2 = 0), dz a complex one. The
// analytic shape and lattice functions are templates over (complex, real) = (cd, double) or
// (dz, dd): instantiated with duals they return the value and its derivative along the direction the
// inputs' d parts hold. A holomorphic f acts as f(v) + e f'(v) d; conj, Re and Im act on both parts
// (the direction is real). Where the plain formula cancels at small arguments (sinc, the sphere and
// cylinder kernels, the Dirichlet kernel near its peaks) the derivative is taken from a series there.
#pragma once

#include "hgx_internal.hpp"

namespace hgx {

// the double versions stay visible beside the dual overloads below
using ::exp;
using ::j0;
using ::j1;
using ::tan;

struct dd {
  double v, d;
};
struct dz {
  cd v, d;
};

// ----- real duals -----
__device__ inline dd operator+(dd a, dd b) { return dd{a.v + b.v, a.d + b.d}; }
__device__ inline dd operator-(dd a, dd b) { return dd{a.v - b.v, a.d - b.d}; }
__device__ inline dd operator-(dd a) { return dd{-a.v, -a.d}; }
__device__ inline dd operator*(dd a, dd b) { return dd{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline dd operator*(double s, dd a) { return dd{s * a.v, s * a.d}; }
__device__ inline dd operator*(dd a, double s) { return dd{s * a.v, s * a.d}; }
__device__ inline dd operator+(dd a, double s) { return dd{a.v + s, a.d}; }
__device__ inline dd operator+(double s, dd a) { return dd{a.v + s, a.d}; }
__device__ inline dd operator-(dd a, double s) { return dd{a.v - s, a.d}; }
__device__ inline dd operator-(double s, dd a) { return dd{s - a.v, -a.d}; }
__device__ inline dd operator/(dd a, dd b) {
  const double q = a.v / b.v;
  return dd{q, (a.d - q * b.d) / b.v};
}
__device__ inline dd operator/(double s, dd b) {
  const double q = s / b.v;
  return dd{q, -q * b.d / b.v};
}
__device__ inline dd operator/(dd a, double s) { return dd{a.v / s, a.d / s}; }
__device__ inline dd exp(dd a) {
  const double e = ::exp(a.v);
  return dd{e, e * a.d};
}
__device__ inline dd tan(dd a) {
  const double t = ::tan(a.v);
  return dd{t, (1.0 + t * t) * a.d};
}
// J1 and its derivative J0 - J1/x (1/2 at 0)
__device__ inline dd j1(dd a) {
  const double j = ::j1(a.v);
  const double dj = a.v == 0.0 ? 0.5 : ::j0(a.v) - j / a.v;
  return dd{j, dj * a.d};
}

// ----- complex duals -----
__device__ inline dz operator+(dz a, dz b) { return dz{a.v + b.v, a.d + b.d}; }
__device__ inline dz operator-(dz a, dz b) { return dz{a.v - b.v, a.d - b.d}; }
__device__ inline dz operator-(dz a) { return dz{-a.v, -a.d}; }
__device__ inline dz operator*(dz a, dz b) { return dz{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline dz operator*(double s, dz a) { return dz{s * a.v, s * a.d}; }
__device__ inline dz operator*(dz a, double s) { return dz{s * a.v, s * a.d}; }
__device__ inline dz operator*(dd s, dz a) { return dz{s.v * a.v, s.d * a.v + s.v * a.d}; }
__device__ inline dz operator*(dz a, dd s) { return s * a; }
__device__ inline dz operator*(cd s, dz a) { return dz{s * a.v, s * a.d}; }
__device__ inline dz operator*(dz a, cd s) { return dz{a.v * s, a.d * s}; }
__device__ inline dd norm2(dz a) { return dd{norm2(a.v), 2.0 * (a.v.r * a.d.r + a.v.i * a.d.i)}; }
__device__ inline dz conj(dz a) { return dz{conj(a.v), conj(a.d)}; }
__device__ inline dz mul_i(dz a) { return dz{mul_i(a.v), mul_i(a.d)}; }
__device__ inline dz mul_mi(dz a) { return dz{mul_mi(a.v), mul_mi(a.d)}; }
__device__ inline dz cdiv(dz a, dz b) {
  const cd q = cdiv(a.v, b.v);
  return dz{q, cdiv(a.d - q * b.d, b.v)};
}
__device__ inline double cabs_d(dz z) { return cabs_d(z.v); }  // branch decisions on the value
__device__ inline dz cexpi_d(dz t) {
  const cd e = cexpi_d(t.v);
  return dz{e, mul_i(e * t.d)};
}
__device__ inline dz cexp_d(dz z) {
  const cd e = cexp_d(z.v);
  return dz{e, e * z.d};
}
__device__ inline dz csin_d(dz z) { return dz{csin_d(z.v), ccos_d(z.v) * z.d}; }
__device__ inline dz ccos_d(dz z) { return dz{ccos_d(z.v), -(csin_d(z.v) * z.d)}; }
__device__ inline dz csqrt_d(dz z) {
  const cd s = csqrt_d(z.v);
  if (s.r == 0.0 && s.i == 0.0) return dz{s, cd{0.0, 0.0}};
  return dz{s, cdiv(z.d, 2.0 * s)};
}

// ----- accessors shared by both instantiations -----
__device__ inline double re(cd z) { return z.r; }
__device__ inline double im(cd z) { return z.i; }
__device__ inline dd re(dz z) { return dd{z.v.r, z.d.r}; }
__device__ inline dd im(dz z) { return dd{z.v.i, z.d.i}; }
__device__ inline double val(double x) { return x; }
__device__ inline double val(dd x) { return x.v; }
__device__ inline cd val(cd z) { return z; }
__device__ inline cd val(dz z) { return z.v; }
__device__ inline cd mkc(double r, double i) { return cd{r, i}; }
__device__ inline dz mkc(dd r, dd i) { return dz{cd{r.v, i.v}, cd{r.d, i.d}}; }
// a complex constant of type C
template <class C>
__device__ inline C cconst(double r, double i = 0.0);
template <>
__device__ inline cd cconst<cd>(double r, double i) {
  return cd{r, i};
}
template <>
__device__ inline dz cconst<dz>(double r, double i) {
  return dz{cd{r, i}, cd{0.0, 0.0}};
}

// ----- derivatives of the kernels that cancel at small arguments -----

// sin(x)/x: value as the plain sinc (free or member form), derivative (cos x - sinc x)/x, or
// -x/3 + x^3/30 - x^5/840 + x^7/45360 for |x| < 0.1 (truncation below 3e-17).
__device__ inline cd dsinc_dx(cd x, cd s) {
  if (cabs_d(x) < 0.1) {
    const cd x2 = x * x;
    return x * (cd{-1.0 / 3.0, 0.0} + x2 * (cd{1.0 / 30.0, 0.0} + x2 * (cd{-1.0 / 840.0, 0.0} + x2 * (1.0 / 45360.0))));
  }
  return cdiv(ccos_d(x) - s, x);
}

// (sin x - x cos x) / x^3: derivative (3 x cos x + (x^2 - 3) sin x) / x^4, or
// -x/15 + x^3/210 - x^5/7560 + x^7/498960 for |x| < 0.1.
__device__ inline cd dsph_dx(cd x) {
  if (cabs_d(x) < 0.1) {
    const cd x2 = x * x;
    return x * (cd{-1.0 / 15.0, 0.0} + x2 * (cd{1.0 / 210.0, 0.0} + x2 * (cd{-1.0 / 7560.0, 0.0} + x2 * (1.0 / 498960.0))));
  }
  const cd x2 = x * x;
  return cdiv(3.0 * (x * ccos_d(x)) + (x2 - cd{3.0, 0.0}) * csin_d(x), x2 * x2);
}

// J1(x)/x for real x: derivative (x J0 - 2 J1)/x^2, or -x/8 + x^3/96 - x^5/3072 + x^7/184320 for
// |x| < 0.1.
__device__ inline double dj1x_dx(double x) {
  if (fabs(x) < 0.1) {
    const double x2 = x * x;
    return x * (-1.0 / 8.0 + x2 * (1.0 / 96.0 + x2 * (-1.0 / 3072.0 + x2 * (1.0 / 184320.0))));
  }
  return (x * ::j0(x) - 2.0 * ::j1(x)) / (x * x);
}

// d/dh of sin(N h)/sin(h): 4th-order series N (-2 a h + 4 b h^3), a = (N^2-1)/6,
// b = (3N^4 - 10N^2 + 7)/360, for |N h| < 1e-2; otherwise the quotient rule (its cancellation error,
// about 3 eps/(N h)^2, stays below 3e-12 there).
__device__ inline cd ddirichlet_dh(cd h, double N) {
  if (cabs_d(N * h) < 1e-2) {
    const double a = (N * N - 1.0) / 6.0, b = (3.0 * N * N * N * N - 10.0 * N * N + 7.0) / 360.0;
    return N * ((-2.0 * a) * h + (4.0 * b) * (h * h * h));
  }
  const cd sh = csin_d(h), ch = ccos_d(h), sn = csin_d(N * h), cn = ccos_d(N * h);
  return cdiv(N * (cn * sh) - sn * ch, sh * sh);
}

}  // namespace hgx
