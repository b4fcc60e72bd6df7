// analytic_shapes.hpp -- FP64 analytic form factors (the CPU formulas of src/ff/ff_ana_<shape>.cpp)
// as device functions of the rotated q-vector, shared by the per-shape kernels and the fused
// lattice-intensity kernels.
#pragma once

#include "hgx_internal.hpp"

namespace hgx {

constexpr double kPi = 3.141592653589793;  // PI_ (include/common/constants.hpp:29)

// Distribution nodes passed by value (kernel arguments live in the constant/scalar path).
struct Nodes {
  int n[4];
  double val[4][HGX_MAX_NODES];
  double wt[4][HGX_MAX_NODES];
  double tan_tau, sin_eta, cos_eta;
  int flags;
};


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

template <int SHAPE>
__device__ cd shape_ff(const cd mq[3], const Nodes& P);

// sphere, src/ff/ff_ana_sphere.cpp:33-45 and :98-110
template <>
__device__ cd shape_ff<HGX_SHAPE_SPHERE>(const cd mq[3], const Nodes& P) {
  cd qval = csqrt_d(mq[0] * mq[0] + mq[1] * mq[1] + mq[2] * mq[2]);
  cd acc = cd{0.0, 0.0};
  if (cabs_d(qval) < 2.2250738585072014e-308) return acc;
  for (int i = 0; i < P.n[0]; ++i) {
    double R = P.val[0][i];
    cd qR = qval * R;
    cd qR3 = qR * qR * qR;
    cd t1 = (1.0 / norm2(qR3)) * conj(qR3);
    cd c0 = (csin_d(qR) - qR * ccos_d(qR)) * t1;
    cd c1 = cexpi_d(mq[2] * R);
    acc = acc + (P.wt[0][i] * (4 * kPi * R * R * R)) * (c0 * c1);
  }
  return acc;
}

// cylinder, src/ff/ff_ana_cylinder.cpp:34-42 and :103-112; J1 of the real part
// (src/numerics/numeric_utils.cpp:69-71)
template <>
__device__ cd shape_ff<HGX_SHAPE_CYLINDER>(const cd mq[3], const Nodes& P) {
  cd qpar = csqrt_d(mq[0] * mq[0] + mq[1] * mq[1]);
  cd acc = cd{0.0, 0.0};
  for (int i = 0; i < P.n[0]; ++i) {
    double R = P.val[0][i];
    cd t1 = qpar * R;
    cd bess = cd{0.0, 0.0};
    double nt = norm2(t1);
    if (nt > 1.0e-13) bess = (j1(t1.r) / nt) * conj(t1);
    for (int k = 0; k < P.n[1]; ++k) {
      double H = P.val[1][k];
      cd hz = (0.5 * H) * mq[2];
      cd v = (2. * kPi * R * R * H) * (sinc_free(hz) * cexpi_d(hz) * bess);
      acc = acc + (P.wt[0][i] * P.wt[1][k]) * v;
    }
  }
  return acc;
}

// box, src/ff/ff_ana_box.cpp:35-41 and :111-121
template <>
__device__ cd shape_ff<HGX_SHAPE_BOX>(const cd mq[3], const Nodes& P) {
  cd acc = cd{0.0, 0.0};
  for (int iz = 0; iz < P.n[2]; ++iz) {
    double H = P.val[2][iz];
    cd hz = (0.5 * H) * mq[2];
    cd ez = cexpi_d(hz) * sinc_free(hz);
    for (int iy = 0; iy < P.n[1]; ++iy) {
      double W = P.val[1][iy];
      cd sy = sinc_free((0.5 * W) * mq[1]);
      for (int ix = 0; ix < P.n[0]; ++ix) {
        double L = P.val[0][ix];
        cd sx = sinc_free((0.5 * L) * mq[0]);
        double w = P.wt[0][ix] * P.wt[1][iy] * P.wt[2][iz];
        acc = acc + (w * (L * W * H)) * (ez * (sx * sy));
      }
    }
  }
  return acc;
}

// cube, src/ff/ff_ana_cube.cpp:35-42 and :109-113
template <>
__device__ cd shape_ff<HGX_SHAPE_CUBE>(const cd mq[3], const Nodes& P) {
  cd acc = cd{0.0, 0.0};
  for (int i = 0; i < P.n[0]; ++i) {
    double L = P.val[0][i];
    cd hz = (0.5 * L) * mq[2];
    cd s = sinc_free((0.5 * L) * mq[0]) * sinc_free((0.5 * L) * mq[1]) * sinc_free(hz);
    acc = acc + (P.wt[0][i] * L * L * L) * (cexpi_d(hz) * s);
  }
  return acc;
}

// triangular prism, src/ff/ff_ana_prism.cpp:87-113
template <>
__device__ cd shape_ff<HGX_SHAPE_PRISM3>(const cd mq[3], const Nodes& P) {
  const double sqrt3 = 1.7320508075688772;
  cd mqx = mq[0], mqy = mq[1], mqz = mq[2];
  cd temp1 = mqx * (mqx * mqx - 3.0 * (mqy * mqy));
  cd temp2 = mqz + P.tan_tau * (mqx * P.sin_eta + mqy * P.cos_eta);
  cd acc = cd{0.0, 0.0};
  for (int il = 0; il < P.n[0]; ++il) {
    double l = P.val[0][il];
    cd temp4 = mqx * cexpi_d((l * sqrt3) * mqy);
    cd temp5 = mqx * ccos_d(l * mqx);
    cd temp6 = mul_i(sqrt3 * (mqy * csin_d(l * mqx)));
    cd temp3 = mqy * (l / sqrt3);
    cd temp9 = (2.0 * sqrt3) * cexp_d(cd{temp3.i, -temp3.r});
    cd a = cdiv(temp9, temp1) * (temp4 - temp5 - temp6);
    for (int ih = 0; ih < P.n[1]; ++ih)
      acc = acc + (P.wt[0][il] * P.wt[1][ih]) * (a * fq_inv(temp2, P.val[1][ih]));
  }
  return acc;
}

// hexagonal prism, src/ff/ff_ana_prism6.cpp:81-101
template <>
__device__ cd shape_ff<HGX_SHAPE_PRISM6>(const cd mq[3], const Nodes& P) {
  const double sqrt3 = 1.7320508075688772;
  cd qm = P.tan_tau * (mq[0] * P.sin_eta + mq[1] * P.cos_eta);
  cd temp1 = cdiv(cd{4.0 * sqrt3, 0.0}, 3.0 * (mq[1] * mq[1]) - mq[0] * mq[0]);
  cd acc = cd{0.0, 0.0};
  for (int ih = 0; ih < P.n[1]; ++ih) {
    cd temp6 = fq_inv(mq[2] + qm, P.val[1][ih]);
    for (int il = 0; il < P.n[0]; ++il) {
      double l = P.val[0][il];
      cd rmqx = (l / sqrt3) * mq[0];
      cd rmqy = l * mq[1];
      cd temp2 = rmqy * rmqy * sinc_member(rmqx) * sinc_member(rmqy);
      cd temp3 = ccos_d(2.0 * rmqx);
      cd temp4 = ccos_d(rmqy) * ccos_d(rmqx);
      cd temp5 = temp1 * (temp2 + temp3 - temp4);
      acc = acc + (P.wt[0][il] * P.wt[1][ih]) * (temp5 * temp6);
    }
  }
  return acc;
}

// truncated pyramid, src/ff/ff_ana_pyramid.cpp:35-74. By default the exponentials are
// exp(+i q1 H), exp(-i q2 H) (published form, the pyramid's volume integral); compat flips them as
// the reference does (:47-60) without flipping the +-i factors of k2, k4.
__device__ cd pyramid_one(cd qx, cd qy, cd qz, double length, double width, double height, double angle,
                          bool compat) {
  double a = angle * kPi / 180.;
  double tan_a = tan(a);
  if ((2 * height / length >= tan_a) || (2 * height / width >= tan_a)) return cd{0.0, 0.0};
  cd tmp = qx * qy;
  if (cabs_d(tmp) < 1.0E-20) return cd{0.0, 0.0};
  cd q1 = 0.5 * ((1.0 / tan_a) * (qx - qy) + qz);
  cd q2 = 0.5 * ((1.0 / tan_a) * (qx - qy) - qz);
  cd q3 = 0.5 * ((1.0 / tan_a) * (qx + qy) + qz);
  cd q4 = 0.5 * ((1.0 / tan_a) * (qx + qy) - qz);
  const double e1 = compat ? -height : height;
  cd a1 = sinc_free(q1 * height) * cexpi_d(e1 * q1);
  cd a2 = sinc_free(q2 * height) * cexpi_d(-e1 * q2);
  cd a3 = sinc_free(q3 * height) * cexpi_d(e1 * q3);
  cd a4 = sinc_free(q4 * height) * cexpi_d(-e1 * q4);
  cd k1 = a1 + a2;
  cd k2 = mul_mi(a1) + mul_i(a2);
  cd k3 = a3 + a4;
  cd k4 = mul_mi(a3) + mul_i(a4);
  cd am = 0.5 * (qx * length - qy * width);
  cd ap = 0.5 * (qx * length + qy * width);
  cd t1 = k1 * ccos_d(am);
  cd t2 = k2 * csin_d(am);
  cd t3 = k3 * ccos_d(ap);
  cd t4 = k4 * csin_d(ap);
  return height * cdiv(t1 + t2 - t3 - t4, tmp);
}

template <>
__device__ cd shape_ff<HGX_SHAPE_PYRAMID>(const cd mq[3], const Nodes& P) {
  cd acc = cd{0.0, 0.0};
  for (int ix = 0; ix < P.n[0]; ++ix)
    for (int iy = 0; iy < P.n[1]; ++iy)
      for (int ih = 0; ih < P.n[2]; ++ih)
        for (int ib = 0; ib < P.n[3]; ++ib) {
          double prob = P.wt[0][ix] * P.wt[1][iy] * P.wt[2][ih] * P.wt[3][ib];
          acc = acc + prob * pyramid_one(mq[0], mq[1], mq[2], P.val[0][ix], P.val[1][iy], P.val[2][ih], P.val[3][ib],
                                         (P.flags & HGX_SHAPE_COMPAT_PYRAMID) != 0);
        }
  return acc;
}

}  // namespace hgx
