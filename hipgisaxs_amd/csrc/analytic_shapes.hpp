// analytic_shapes.hpp -- FP64 analytic form factors (the CPU formulas of src/ff/ff_ana_<shape>.cpp)
// as device functions of the rotated q-vector, shared by the per-shape kernels and the fused
// lattice-intensity kernels. Each formula is a template over the (complex, real) types: (cd, double)
// gives the value, (dz, dd) the value with its directional derivative (dual.hpp).
#pragma once

#include <type_traits>

#include "dual.hpp"

namespace hgx {

constexpr double kPi = 3.141592653589793;  // PI_ (include/common/constants.hpp:29)

// Distribution nodes passed by value (kernel arguments live in the constant/scalar path); R = dd
// carries the nodes' tangents.
template <class R>
struct NodesT {
  int n[4];
  R val[4][HGX_MAX_NODES];
  R wt[4][HGX_MAX_NODES];
  R tan_tau, sin_eta, cos_eta;
  int flags;
};
using Nodes = NodesT<double>;


// free sinc of include/numerics/numeric_utils.hpp:54-64
__device__ inline cd sinc_free(cd x) {
  if (cabs_d(x) > 1.0e-14) return cdiv(csin_d(x), x);
  cd x2 = x * x;
  return cd{1.0, 0.0} - (1.0 / 6.0) * x2 + (1.0 / 120.0) * (x2 * x2);
}
// member sinc of src/ff/ff_ana.cpp:224-229 (prism6)
__device__ inline cd sinc_member(cd v) {
  if (fabs(v.r) <= 1e-14 && fabs(v.i) <= 1e-14) return cd{1.0, 0.0};
  return cdiv(csin_d(v), v);
}
// fq_inv of src/ff/ff_ana.cpp:190-196
__device__ inline cd fq_inv(cd value, double y) {
  cd half = (0.5 * y) * value;
  cd temp = 2.0 * cdiv(cexpi_d(half) * csin_d(half), value);
  if (cabs_d(temp) <= 1e-14) temp = cd{y, 0.0};
  return temp;
}
// sphere kernel (sin x - x cos x) / x^3, as src/ff/ff_ana_sphere.cpp:33-45 forms it
__device__ inline cd sph_core(cd qR) {
  cd qR3 = qR * qR * qR;
  cd t1 = (1.0 / norm2(qR3)) * conj(qR3);
  return (csin_d(qR) - qR * ccos_d(qR)) * t1;
}
// cylinder kernel J1(Re t) conj(t) / |t|^2 (0 below |t|^2 = 1e-13), src/ff/ff_ana_cylinder.cpp:34-42
__device__ inline cd cyl_core(cd t1) {
  cd bess = cd{0.0, 0.0};
  double nt = norm2(t1);
  if (nt > 1.0e-13) bess = (j1(t1.r) / nt) * conj(t1);
  return bess;
}

// ----- the same kernels on duals: plain values, derivatives free of small-argument cancellation -----
__device__ inline dz sinc_free(dz x) {
  const cd s = sinc_free(x.v);
  return dz{s, dsinc_dx(x.v, s) * x.d};
}
__device__ inline dz sinc_member(dz x) {
  const cd s = sinc_member(x.v);
  return dz{s, dsinc_dx(x.v, s) * x.d};
}
// fq_inv = y exp(i h) sin(h)/h with h = y v / 2 (the branch for |value| <= 1e-14 returns y)
__device__ inline dz fq_inv(dz value, dd y) {
  const cd v = fq_inv(value.v, y.v);
  const cd half = (0.5 * y.v) * value.v;
  if (cabs_d(2.0 * cdiv(cexpi_d(half) * csin_d(half), value.v)) <= 1e-14) return dz{v, cd{y.d, 0.0}};
  const cd s = sinc_member(half), e = cexpi_d(half);
  const cd dh = 0.5 * (y.d * value.v + y.v * value.d);
  // d[y e^{ih} sinc h] = dy e^{ih} sinc h + y e^{ih} (i sinc h + sinc'(h)) dh
  const cd dv = y.d * (e * s) + y.v * (e * ((mul_i(s) + dsinc_dx(half, s)) * dh));
  return dz{v, dv};
}
__device__ inline dz sph_core(dz qR) { return dz{sph_core(qR.v), dsph_dx(qR.v) * qR.d}; }
__device__ inline dz cyl_core(dz t) {
  const cd b = cyl_core(t.v);
  const double nt = norm2(t.v);
  if (!(nt > 1.0e-13)) return dz{b, cd{0.0, 0.0}};
  const double a = t.v.r;
  if (t.v.i == 0.0 && fabs(a) < 0.1) {
    // b = J1(a)/a at real t: d/da by the series, and -i J1(a)/a^2 along Im t
    return dz{b, cd{dj1x_dx(a) * t.d.r, 0.0} + cd{0.0, -::j1(a) / (a * a) * t.d.i}};
  }
  // J1(a) conj(t) / |t|^2
  const double j = ::j1(a), dj = a == 0.0 ? 0.5 : ::j0(a) - j / a;
  const double dnt = 2.0 * (t.v.r * t.d.r + t.v.i * t.d.i);
  const cd db = ((dj * t.d.r) / nt) * conj(t.v) + (j / nt) * conj(t.d) - (j * dnt / (nt * nt)) * conj(t.v);
  return dz{b, db};
}

template <int SHAPE>
struct ShapeFF;

// sphere, src/ff/ff_ana_sphere.cpp:33-45 and :98-110
template <>
struct ShapeFF<HGX_SHAPE_SPHERE> {
  // the reference GPU's cap of height H (ff_ana_sphere_gpu.cu:53-86, 100-108): trapezoid over
  // kCapSlices disks, each 0 where |q_par| < 1e-13 or r Re q_par < 1e-13
  static constexpr int kCapSlices = 512;
  __device__ static cd cap(const cd mq[3], double Rr, double H) {
    const cd qp = csqrt_d(mq[0] * mq[0] + mq[1] * mq[1]);
    if (cabs_d(qp) < 1e-13) return cd{0.0, 0.0};
    const double cen = H - Rr, dh = H / (double)(kCapSlices - 1);
    cd out{0.0, 0.0};
    for (int s = 0; s < kCapSlices; ++s) {
      const double z = s * dh - cen;
      // the last slice can pass the pole by a rounding step (511 (H / 511) > H): an empty disk there,
      // where the reference takes the root of a negative number
      const double rz = sqrt(fmax(Rr * Rr - z * z, 0.0));
      const double t2 = rz * qp.r;
      if (t2 < 1e-13) continue;
      const double disk = 2 * kPi * rz * rz * (::j1(t2) / t2);
      out = out + ((s == 0 || s == kCapSlices - 1) ? disk : 2.0 * disk) * cexpi_d(mq[2] * z);
    }
    return (0.5 * dh) * out;
  }
  template <class C, class R>
  __device__ static C eval(const C mq[3], const NodesT<R>& P) {
    if constexpr (std::is_same<R, double>::value) {
      if ((P.flags & HGX_SHAPE_TRUNCATED_SPHERE) && P.n[1] > 0) {
        C acc = cconst<C>(0.0);
        for (int i = 0; i < P.n[0]; ++i)
          for (int k = 0; k < P.n[1]; ++k) {
            const double Rr = P.val[0][i], H = P.val[1][k];
            acc = acc + (P.wt[1][k] * P.wt[0][i]) * (cexpi_d(mq[2] * (H - Rr)) * cap(mq, Rr, H));
          }
        return acc;
      }
    }
    return plain(mq, P);
  }
  // the full sphere (any flags ignored)
  template <class C, class R>
  __device__ static C plain(const C mq[3], const NodesT<R>& P) {
    C qval = csqrt_d(mq[0] * mq[0] + mq[1] * mq[1] + mq[2] * mq[2]);
    C acc = cconst<C>(0.0);
    if (cabs_d(qval) < 2.2250738585072014e-308) return acc;
    for (int i = 0; i < P.n[0]; ++i) {
      R Rr = P.val[0][i];
      C qR = qval * Rr;
      C c0 = sph_core(qR);
      C c1 = cexpi_d(mq[2] * Rr);
      acc = acc + (P.wt[0][i] * (4 * kPi * Rr * Rr * Rr)) * (c0 * c1);
    }
    return acc;
  }
};

// An internal shape id for spheres the host knows are not truncated: the same formula without the
// cap's code path (its FP64 Bessel function over 512 slices), so that kernels instantiated for it keep
// their registers and need no scratch.
constexpr int kShapePlainSphere = 8;

template <>
struct ShapeFF<kShapePlainSphere> {
  template <class C, class R>
  __device__ static C eval(const C mq[3], const NodesT<R>& P) {
    return ShapeFF<HGX_SHAPE_SPHERE>::plain(mq, P);
  }
};

// cylinder, src/ff/ff_ana_cylinder.cpp:34-42 and :103-112; J1 of the real part
// (src/numerics/numeric_utils.cpp:69-71)
template <>
struct ShapeFF<HGX_SHAPE_CYLINDER> {
  template <class C, class R>
  __device__ static C eval(const C mq[3], const NodesT<R>& P) {
    C qpar = csqrt_d(mq[0] * mq[0] + mq[1] * mq[1]);
    C acc = cconst<C>(0.0);
    for (int i = 0; i < P.n[0]; ++i) {
      R Rr = P.val[0][i];
      C bess = cyl_core(qpar * Rr);
      for (int k = 0; k < P.n[1]; ++k) {
        R H = P.val[1][k];
        C hz = (0.5 * H) * mq[2];
        C v = (2. * kPi * Rr * Rr * H) * (sinc_free(hz) * cexpi_d(hz) * bess);
        acc = acc + (P.wt[0][i] * P.wt[1][k]) * v;
      }
    }
    return acc;
  }
};

// box, src/ff/ff_ana_box.cpp:35-41 and :111-121
template <>
struct ShapeFF<HGX_SHAPE_BOX> {
  template <class C, class R>
  __device__ static C eval(const C mq[3], const NodesT<R>& P) {
    C acc = cconst<C>(0.0);
    for (int iz = 0; iz < P.n[2]; ++iz) {
      R H = P.val[2][iz];
      C hz = (0.5 * H) * mq[2];
      C ez = cexpi_d(hz) * sinc_free(hz);
      for (int iy = 0; iy < P.n[1]; ++iy) {
        R W = P.val[1][iy];
        C sy = sinc_free((0.5 * W) * mq[1]);
        for (int ix = 0; ix < P.n[0]; ++ix) {
          R L = P.val[0][ix];
          C sx = sinc_free((0.5 * L) * mq[0]);
          R w = P.wt[0][ix] * P.wt[1][iy] * P.wt[2][iz];
          acc = acc + (w * (L * W * H)) * (ez * (sx * sy));
        }
      }
    }
    return acc;
  }
};

// cube, src/ff/ff_ana_cube.cpp:35-42 and :109-113
template <>
struct ShapeFF<HGX_SHAPE_CUBE> {
  template <class C, class R>
  __device__ static C eval(const C mq[3], const NodesT<R>& P) {
    C acc = cconst<C>(0.0);
    for (int i = 0; i < P.n[0]; ++i) {
      R L = P.val[0][i];
      C hz = (0.5 * L) * mq[2];
      C s = sinc_free((0.5 * L) * mq[0]) * sinc_free((0.5 * L) * mq[1]) * sinc_free(hz);
      acc = acc + (P.wt[0][i] * L * L * L) * (cexpi_d(hz) * s);
    }
    return acc;
  }
};

// triangular prism, src/ff/ff_ana_prism.cpp:87-113
template <>
struct ShapeFF<HGX_SHAPE_PRISM3> {
  template <class C, class R>
  __device__ static C eval(const C mq[3], const NodesT<R>& P) {
    const double sqrt3 = 1.7320508075688772;
    C mqx = mq[0], mqy = mq[1], mqz = mq[2];
    C temp1 = mqx * (mqx * mqx - 3.0 * (mqy * mqy));
    C temp2 = mqz + P.tan_tau * (mqx * P.sin_eta + mqy * P.cos_eta);
    C acc = cconst<C>(0.0);
    for (int il = 0; il < P.n[0]; ++il) {
      R l = P.val[0][il];
      C temp4 = mqx * cexpi_d((l * sqrt3) * mqy);
      C temp5 = mqx * ccos_d(l * mqx);
      C temp6 = mul_i(sqrt3 * (mqy * csin_d(l * mqx)));
      C temp3 = mqy * (l / sqrt3);
      C temp9 = (2.0 * sqrt3) * cexp_d(mul_mi(temp3));
      C a = cdiv(temp9, temp1) * (temp4 - temp5 - temp6);
      for (int ih = 0; ih < P.n[1]; ++ih)
        acc = acc + (P.wt[0][il] * P.wt[1][ih]) * (a * fq_inv(temp2, P.val[1][ih]));
    }
    return acc;
  }
};

// hexagonal prism, src/ff/ff_ana_prism6.cpp:81-101
template <>
struct ShapeFF<HGX_SHAPE_PRISM6> {
  template <class C, class R>
  __device__ static C eval(const C mq[3], const NodesT<R>& P) {
    const double sqrt3 = 1.7320508075688772;
    C qm = P.tan_tau * (mq[0] * P.sin_eta + mq[1] * P.cos_eta);
    C temp1 = cdiv(cconst<C>(4.0 * sqrt3, 0.0), 3.0 * (mq[1] * mq[1]) - mq[0] * mq[0]);
    C acc = cconst<C>(0.0);
    for (int ih = 0; ih < P.n[1]; ++ih) {
      C temp6 = fq_inv(mq[2] + qm, P.val[1][ih]);
      for (int il = 0; il < P.n[0]; ++il) {
        R l = P.val[0][il];
        C rmqx = (l / sqrt3) * mq[0];
        C rmqy = l * mq[1];
        C temp2 = rmqy * rmqy * sinc_member(rmqx) * sinc_member(rmqy);
        C temp3 = ccos_d(2.0 * rmqx);
        C temp4 = ccos_d(rmqy) * ccos_d(rmqx);
        C temp5 = temp1 * (temp2 + temp3 - temp4);
        acc = acc + (P.wt[0][il] * P.wt[1][ih]) * (temp5 * temp6);
      }
    }
    return acc;
  }
};

// truncated pyramid, src/ff/ff_ana_pyramid.cpp:35-74. By default the exponentials are
// exp(+i q1 H), exp(-i q2 H) (published form, the pyramid's volume integral); compat flips them as
// the reference does (:47-60) without flipping the +-i factors of k2, k4.
template <class C, class R>
__device__ C pyramid_one(C qx, C qy, C qz, R length, R width, R height, R angle, bool compat) {
  R a = angle * kPi / 180.;
  R tan_a = tan(a);
  if ((2 * val(height) / val(length) >= val(tan_a)) || (2 * val(height) / val(width) >= val(tan_a)))
    return cconst<C>(0.0);
  C tmp = qx * qy;
  if (cabs_d(tmp) < 1.0E-20) return cconst<C>(0.0);
  C q1 = 0.5 * ((1.0 / tan_a) * (qx - qy) + qz);
  C q2 = 0.5 * ((1.0 / tan_a) * (qx - qy) - qz);
  C q3 = 0.5 * ((1.0 / tan_a) * (qx + qy) + qz);
  C q4 = 0.5 * ((1.0 / tan_a) * (qx + qy) - qz);
  const R e1 = compat ? -height : height;
  C a1 = sinc_free(q1 * height) * cexpi_d(e1 * q1);
  C a2 = sinc_free(q2 * height) * cexpi_d(-e1 * q2);
  C a3 = sinc_free(q3 * height) * cexpi_d(e1 * q3);
  C a4 = sinc_free(q4 * height) * cexpi_d(-e1 * q4);
  C k1 = a1 + a2;
  C k2 = mul_mi(a1) + mul_i(a2);
  C k3 = a3 + a4;
  C k4 = mul_mi(a3) + mul_i(a4);
  C am = 0.5 * (qx * length - qy * width);
  C ap = 0.5 * (qx * length + qy * width);
  C t1 = k1 * ccos_d(am);
  C t2 = k2 * csin_d(am);
  C t3 = k3 * ccos_d(ap);
  C t4 = k4 * csin_d(ap);
  return height * cdiv(t1 + t2 - t3 - t4, tmp);
}

template <>
struct ShapeFF<HGX_SHAPE_PYRAMID> {
  template <class C, class R>
  __device__ static C eval(const C mq[3], const NodesT<R>& P) {
    C acc = cconst<C>(0.0);
    for (int ix = 0; ix < P.n[0]; ++ix)
      for (int iy = 0; iy < P.n[1]; ++iy)
        for (int ih = 0; ih < P.n[2]; ++ih)
          for (int ib = 0; ib < P.n[3]; ++ib) {
            R prob = P.wt[0][ix] * P.wt[1][iy] * P.wt[2][ih] * P.wt[3][ib];
            acc = acc + prob * pyramid_one(mq[0], mq[1], mq[2], P.val[0][ix], P.val[1][iy], P.val[2][ih],
                                           P.val[3][ib], (P.flags & HGX_SHAPE_COMPAT_PYRAMID) != 0);
          }
    return acc;
  }
};

// F of one shape at the rotated q: the plain FP64 value ...
template <int SHAPE>
__device__ inline cd shape_ff(const cd mq[3], const Nodes& P) {
  return ShapeFF<SHAPE>::template eval<cd, double>(mq, P);
}
// ... and the value with its derivative along the tangents of q and of the nodes
template <int SHAPE>
__device__ inline dz shape_ff(const dz mq[3], const NodesT<dd>& P) {
  return ShapeFF<SHAPE>::template eval<dz, dd>(mq, P);
}

}  // namespace hgx
