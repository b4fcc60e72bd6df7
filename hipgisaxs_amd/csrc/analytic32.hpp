// analytic32.hpp -- FP32C analytic form factors (real q): the CPU formulas of the reference
// (src/ff/ff_ana_<shape>.cpp) with every phase formed and reduced to turns in FP64, the hardware
// turn-based sin/cos in FP32 (Taylor series at small arguments) and node sums in FP32.
//
// The prism and pyramid closed forms divide by polynomials in q that vanish on lines where the
// numerator vanishes too (prism3: qx^2 = 3 qy^2; prism6: 3 qy^2 = qx^2; pyramid: qx qy = 0) and at
// small |q|. Near those lines the FP32 numerator would lose its leading digits, so ff_fp32c falls
// back, per q-point, to the FP64 formula where the cancellation exceeds a fixed bound (the FP32C
// principle: FP64 where cancellation needs it). The bounds keep the FP32 amplification below about 20.
#pragma once

#include <cmath>

#include "analytic_shapes.hpp"
#include "fp32_math.hpp"
#include "lattice_sf.hpp"

namespace hgx {

// x (radians, FP64) reduced to turns in (-1/2, 1/2]
__device__ inline float turns_of(double x) {
  double u = x * (1.0 / (2.0 * kPiD));
  return (float)(u - rint(u));
}

// sin(u)/u for u = 0.5 * m * L, the phase reduced in FP64
__device__ inline float sinc32(double m, double L) {
  double u = 0.5 * m * L;
  float uf = (float)u;
  if (fabsf(uf) < 0.1f) {
    float u2 = uf * uf;
    return 1.0f - u2 * (1.0f / 6.0f) * (1.0f - u2 * (1.0f / 20.0f));
  }
  return sin_turn(turns_of(u)) / uf;
}

template <int SHAPE>
__device__ cf ff32(const double m[3], const Nodes& P);

// sphere (src/ff/ff_ana_sphere.cpp:33-45): sum_k w_k 4 pi R^3 (sin x - x cos x)/x^3 e^{i m_z R}
template <>
__device__ cf ff32<HGX_SHAPE_SPHERE>(const double m[3], const Nodes& P) {
  const double q64 = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  if (q64 == 0.0) return cf{0.0f, 0.0f};
  const float qv = (float)q64;
  cf acc = cf{0.0f, 0.0f};
  for (int i = 0; i < P.n[0]; ++i) {
    const double R = P.val[0][i];
    const float x = qv * (float)R;
    float t;
    if (x < 0.6f) {
      float x2 = x * x;
      t = x2 * x * ((1.0f / 3.0f) - x2 * ((1.0f / 30.0f) - x2 * ((1.0f / 840.0f) - x2 * (1.0f / 45360.0f))));
    } else {
      const float tr = turns_of(q64 * R);
      t = sin_turn(tr) - x * cos_turn(tr);
    }
    acc = acc + ((float)(P.wt[0][i] * 4.0 * kPiD) * t) * expi_turn(turns_of(m[2] * R));
  }
  return (1.0f / (qv * qv * qv)) * acc;
}

// cylinder (src/ff/ff_ana_cylinder.cpp:34-42), node double sum factorised
template <>
__device__ cf ff32<HGX_SHAPE_CYLINDER>(const double m[3], const Nodes& P) {
  const float qp = (float)sqrt(m[0] * m[0] + m[1] * m[1]);
  float br = 0.0f;
  for (int i = 0; i < P.n[0]; ++i) {
    const float R = (float)P.val[0][i];
    if (qp * R * qp * R > 1.0e-13f) br += (float)(P.wt[0][i] * 2.0 * kPiD) * R * R * j1_over_x(qp * R);
  }
  cf hz = cf{0.0f, 0.0f};
  for (int k = 0; k < P.n[1]; ++k) {
    const double H = P.val[1][k];
    hz = hz + ((float)(P.wt[1][k] * H) * sinc32(m[2], H)) * expi_turn(turns_of(0.5 * H * m[2]));
  }
  return br * hz;
}

// box (src/ff/ff_ana_box.cpp:35-41, :111-121), node triple sum factorised per axis
template <>
__device__ cf ff32<HGX_SHAPE_BOX>(const double m[3], const Nodes& P) {
  float sx = 0.0f, sy = 0.0f;
  for (int i = 0; i < P.n[0]; ++i) sx += (float)(P.wt[0][i] * P.val[0][i]) * sinc32(m[0], P.val[0][i]);
  for (int i = 0; i < P.n[1]; ++i) sy += (float)(P.wt[1][i] * P.val[1][i]) * sinc32(m[1], P.val[1][i]);
  cf hz = cf{0.0f, 0.0f};
  for (int k = 0; k < P.n[2]; ++k) {
    const double H = P.val[2][k];
    hz = hz + ((float)(P.wt[2][k] * H) * sinc32(m[2], H)) * expi_turn(turns_of(0.5 * H * m[2]));
  }
  return (sx * sy) * hz;
}

// cube (src/ff/ff_ana_cube.cpp:35-42, :109-113)
template <>
__device__ cf ff32<HGX_SHAPE_CUBE>(const double m[3], const Nodes& P) {
  cf acc = cf{0.0f, 0.0f};
  for (int i = 0; i < P.n[0]; ++i) {
    const double L = P.val[0][i];
    const float s = sinc32(m[0], L) * sinc32(m[1], L) * sinc32(m[2], L);
    acc = acc + ((float)(P.wt[0][i] * L * L * L) * s) * expi_turn(turns_of(0.5 * L * m[2]));
  }
  return acc;
}

// sin(x)/x of an FP64 argument: the phase reduced in FP64, the sine in FP32 (series below 0.1)
__device__ inline float sincx32(double x) {
  const float xf = (float)x;
  if (fabsf(xf) < 0.1f) {
    const float x2 = xf * xf;
    return 1.0f - x2 * (1.0f / 6.0f) * (1.0f - x2 * (1.0f / 20.0f));
  }
  return sin_turn(turns_of(x)) / xf;
}

// triangular prism (src/ff/ff_ana_prism.cpp:87-113) for real q: with phases p2 = sqrt3 l qy,
// p3 = l qx, p1 = -l qy / sqrt3,
//   F = [sum_l w_l (2 sqrt3 / (qx (qx^2 - 3 qy^2))) e^{i p1} (qx (cos p2 - cos p3) + i (qx sin p2 - sqrt3 qy sin p3))]
//     x [sum_h w_h h sinc(h t2 / 2) e^{i h t2 / 2}],  t2 = qz + tan(tau)(qx sin(eta) + qy cos(eta))
// (the node double sum factorises; fq_inv(v, h) of src/ff/ff_ana.cpp:190-196 is h sinc(h v/2) e^{i h v/2}).
template <>
__device__ cf ff32<HGX_SHAPE_PRISM3>(const double m[3], const Nodes& P) {
  const double s3 = 1.7320508075688772;
  const double qx = m[0], qy = m[1], qz = m[2];
  const double t2 = qz + P.tan_tau * (qx * P.sin_eta + qy * P.cos_eta);
  cf hs = cf{0.0f, 0.0f};
  for (int ih = 0; ih < P.n[1]; ++ih) {
    const double h = P.val[1][ih];
    hs = hs + ((float)(P.wt[1][ih] * h) * sinc32(t2, h)) * expi_turn(turns_of(0.5 * h * t2));
  }
  const float qxf = (float)qx, qyf = (float)qy;
  cf acc = cf{0.0f, 0.0f};
  for (int il = 0; il < P.n[0]; ++il) {
    const double l = P.val[0][il];
    const float t2a = turns_of(s3 * l * qy), t3a = turns_of(l * qx);
    const float c2 = cos_turn(t2a), s2 = sin_turn(t2a), c3 = cos_turn(t3a), sn3 = sin_turn(t3a);
    const cf B = cf{qxf * (c2 - c3), qxf * s2 - (float)s3 * qyf * sn3};
    acc = acc + (float)P.wt[0][il] * (expi_turn(turns_of(-l * qy / s3)) * B);
  }
  const float scale = (float)(2.0 * s3 / (qx * (qx * qx - 3.0 * qy * qy)));
  return (scale * acc) * hs;
}

// hexagonal prism (src/ff/ff_ana_prism6.cpp:81-101) for real q: with rx = l qx / sqrt3, ry = l qy,
//   F = 4 sqrt3 / (3 qy^2 - qx^2) [sum_l w_l (ry^2 sinc(rx) sinc(ry) + cos 2rx - cos ry cos rx)]
//     x [sum_h w_h h sinc(h v / 2) e^{i h v / 2}],  v = qz + tan(tau)(qx sin(eta) + qy cos(eta)).
template <>
__device__ cf ff32<HGX_SHAPE_PRISM6>(const double m[3], const Nodes& P) {
  const double s3 = 1.7320508075688772;
  const double qx = m[0], qy = m[1], qz = m[2];
  const double v = qz + P.tan_tau * (qx * P.sin_eta + qy * P.cos_eta);
  cf hs = cf{0.0f, 0.0f};
  for (int ih = 0; ih < P.n[1]; ++ih) {
    const double h = P.val[1][ih];
    hs = hs + ((float)(P.wt[1][ih] * h) * sinc32(v, h)) * expi_turn(turns_of(0.5 * h * v));
  }
  float acc = 0.0f;
  for (int il = 0; il < P.n[0]; ++il) {
    const double l = P.val[0][il];
    const double rx = l * qx / s3, ry = l * qy;
    const float ryf = (float)ry;
    const float t2 = ryf * ryf * sincx32(rx) * sincx32(ry);
    const float t3 = cos_turn(turns_of(2.0 * rx));
    const float t4 = cos_turn(turns_of(ry)) * cos_turn(turns_of(rx));
    acc += (float)P.wt[0][il] * (t2 + t3 - t4);
  }
  const float scale = (float)(4.0 * s3 / (3.0 * qy * qy - qx * qx));
  return (scale * acc) * hs;
}

// truncated pyramid (src/ff/ff_ana_pyramid.cpp:35-74, the published signs by default as
// pyramid_one) for real q
__device__ inline cf pyramid32_one(double qx, double qy, double qz, double length, double width, double height,
                                   double angle, bool compat) {
  const double a = angle * kPi / 180.;
  const double tan_a = tan(a);
  if ((2 * height / length >= tan_a) || (2 * height / width >= tan_a)) return cf{0.0f, 0.0f};
  const double tmp = qx * qy;
  if (fabs(tmp) < 1.0E-20) return cf{0.0f, 0.0f};
  const double q1 = 0.5 * ((qx - qy) / tan_a + qz), q2 = 0.5 * ((qx - qy) / tan_a - qz);
  const double q3 = 0.5 * ((qx + qy) / tan_a + qz), q4 = 0.5 * ((qx + qy) / tan_a - qz);
  const double e1 = compat ? -height : height;
  const cf a1 = sincx32(q1 * height) * expi_turn(turns_of(e1 * q1));
  const cf a2 = sincx32(q2 * height) * expi_turn(turns_of(-e1 * q2));
  const cf a3 = sincx32(q3 * height) * expi_turn(turns_of(e1 * q3));
  const cf a4 = sincx32(q4 * height) * expi_turn(turns_of(-e1 * q4));
  const cf k1 = a1 + a2, k3 = a3 + a4;
  const cf k2 = cf{a1.i - a2.i, a2.r - a1.r};  // -i a1 + i a2
  const cf k4 = cf{a3.i - a4.i, a4.r - a3.r};
  const float am = turns_of(0.5 * (qx * length - qy * width)), ap = turns_of(0.5 * (qx * length + qy * width));
  const float cm = cos_turn(am), sm = sin_turn(am), cp = cos_turn(ap), sp = sin_turn(ap);
  const cf num = cf{k1.r * cm + k2.r * sm - k3.r * cp - k4.r * sp, k1.i * cm + k2.i * sm - k3.i * cp - k4.i * sp};
  return (float)(height / tmp) * num;
}

template <>
__device__ cf ff32<HGX_SHAPE_PYRAMID>(const double m[3], const Nodes& P) {
  cf acc = cf{0.0f, 0.0f};
  const bool compat = (P.flags & HGX_SHAPE_COMPAT_PYRAMID) != 0;
  for (int ix = 0; ix < P.n[0]; ++ix)
    for (int iy = 0; iy < P.n[1]; ++iy)
      for (int ih = 0; ih < P.n[2]; ++ih)
        for (int ib = 0; ib < P.n[3]; ++ib) {
          const float prob = (float)(P.wt[0][ix] * P.wt[1][iy] * P.wt[2][ih] * P.wt[3][ib]);
          acc = acc + prob * pyramid32_one(m[0], m[1], m[2], P.val[0][ix], P.val[1][iy], P.val[2][ih],
                                           P.val[3][ib], compat);
        }
  return acc;
}

__device__ inline double min_node(const Nodes& P, int d) {
  double v = P.val[d][0];
  for (int i = 1; i < P.n[d]; ++i) v = fmin(v, P.val[d][i]);
  return v;
}

// Whether the FP32 evaluation at m is well conditioned (else ff_fp32c takes the FP64 formula): away
// from small |q| (|q_par| l >= 1) and from the lines where the closed form's denominator vanishes
// (its numerator cancels there by the factor l |distance to the line|, kept >= 0.05; for the pyramid
// |qx| L, |qy| W >= 0.5).
template <int SHAPE>
__device__ inline bool fp32_conditioned(const double m[3], const Nodes& P) {
  if (SHAPE == HGX_SHAPE_PRISM3 || SHAPE == HGX_SHAPE_PRISM6) {
    const double s3 = 1.7320508075688772;
    const double l = min_node(P, 0);
    const double qp = sqrt(m[0] * m[0] + m[1] * m[1]);
    if (qp * l < 1.0) return false;
    const double d = SHAPE == HGX_SHAPE_PRISM3 ? fmin(fabs(m[0] - s3 * m[1]), fabs(m[0] + s3 * m[1]))
                                               : fmin(fabs(m[0] / s3 - m[1]), fabs(m[0] / s3 + m[1]));
    return l * d >= 0.05;
  }
  if (SHAPE == HGX_SHAPE_PYRAMID) return fabs(m[0]) * min_node(P, 0) >= 0.5 && fabs(m[1]) * min_node(P, 1) >= 0.5;
  if (SHAPE == HGX_SHAPE_SPHERE) return !(P.flags & HGX_SHAPE_TRUNCATED_SPHERE);  // the cap: FP64
  return true;
}

// the plain sphere (kShapePlainSphere) is the sphere's FP32 form, always conditioned
template <>
__device__ inline cf ff32<kShapePlainSphere>(const double m[3], const Nodes& P) {
  return ff32<HGX_SHAPE_SPHERE>(m, P);
}

// FP32C form factor at a real rotated q-point, returned in FP64.
template <int SHAPE>
__device__ inline cd ff_fp32c(const double m[3], const Nodes& P) {
  if constexpr (SHAPE == kShapePlainSphere) {
    const cf F = ff32<SHAPE>(m, P);
    return cd{(double)F.r, (double)F.i};
  }
  if (!fp32_conditioned<SHAPE>(m, P)) {
    const cd mq[3] = {cd{m[0], 0.0}, cd{m[1], 0.0}, cd{m[2], 0.0}};
    return shape_ff<SHAPE>(mq, P);
  }
  const cf F = ff32<SHAPE>(m, P);
  return cd{(double)F.r, (double)F.i};
}

// S(m) in FP32 from FP64 lattice phases (the FP32C lattice kernel's form)
__device__ inline cf sf32e(const double m[3], const SFArgs& s) {
  const double xa = m[0] * s.la[0] + m[1] * s.la[1] + m[2] * s.la[2];
  const double xb = m[0] * s.lb[0] + m[1] * s.lb[1] + (s.compat_lb2 ? s.lb[2] * s.lb[2] : m[2] * s.lb[2]);
  const double xc = m[0] * s.lc[0] + m[1] * s.lc[1] + m[2] * s.lc[2];
  float d = dirichlet_turn(turns_of(xa), (float)s.N[0]) * dirichlet_turn(turns_of(xb), (float)s.N[1]) *
            dirichlet_turn(turns_of(xc), (float)s.N[2]);
  const cf ph = expi_turn(turns_of(m[0] * s.center[0] + m[1] * s.center[1] + m[2] * s.center[2]));
  if (s.dw != 0.0) d *= __expf((float)(-s.dw * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2])));
  return d * ph;
}

// the same with the shape chosen at run time (cells of mixed shapes)
__device__ inline cd ff_fp32c_any(int shape, const double m[3], const Nodes& P) {
  switch (shape) {
    case HGX_SHAPE_SPHERE: return ff_fp32c<HGX_SHAPE_SPHERE>(m, P);
    case HGX_SHAPE_CYLINDER: return ff_fp32c<HGX_SHAPE_CYLINDER>(m, P);
    case HGX_SHAPE_BOX: return ff_fp32c<HGX_SHAPE_BOX>(m, P);
    case HGX_SHAPE_CUBE: return ff_fp32c<HGX_SHAPE_CUBE>(m, P);
    case HGX_SHAPE_PRISM3: return ff_fp32c<HGX_SHAPE_PRISM3>(m, P);
    case HGX_SHAPE_PRISM6: return ff_fp32c<HGX_SHAPE_PRISM6>(m, P);
    default: return ff_fp32c<HGX_SHAPE_PYRAMID>(m, P);
  }
}

}  // namespace hgx
