// ff_analytic.hip -- analytic form factors (FP64), one thread per q-point, node loops in registers.
// Formulas are the CPU ones of the reference (src/ff/ff_ana_<shape>.cpp), which the reference's
// GPU kernels do not all follow (src/ff/gpu/ff_ana_prism_gpu.cu:39-57, ff_ana_prism6_gpu.cu:47).
#include <cmath>

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

struct QViewA {
  const double* __restrict__ qx;
  const double* __restrict__ qy;
  const double* __restrict__ qz;
  int64_t npix, nq;
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

template <int SHAPE>
__global__ __launch_bounds__(256) void k_ff_analytic(QViewA g, Rot R, Nodes P, Vec3 tv, int has_tv,
                                                     double* __restrict__ out) {
  int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= g.nq) return;
  int64_t p = q % g.npix;
  cd mq[3];
  rotate_q(R, g.qx[p], g.qy[p], cd{g.qz[2 * q], g.qz[2 * q + 1]}, mq);
  cd F = shape_ff<SHAPE>(mq, P);
  if (has_tv) F = F * cexpi_d(mq[0] * tv.v[0] + mq[1] * tv.v[1] + mq[2] * tv.v[2]);
  out[2 * q] = F.r;
  out[2 * q + 1] = F.i;
}

}  // namespace hgx

using namespace hgx;

static const int kDims[7] = {1, 2, 3, 1, 2, 2, 4};

extern "C" int hgx_ff_analytic(hgx_ctx* h, int shape, const hgx_shape_params* params, int precision,
                               const double rot[9], const double transvec[3], double* ff) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "ff_analytic: no q-grid");
  if (shape < 0 || shape > HGX_SHAPE_PYRAMID) return set_error(c, HGX_ERR_INVALID_ARG, "ff_analytic: bad shape %d", shape);
  if (!params || !ff) return set_error(c, HGX_ERR_INVALID_ARG, "ff_analytic: null argument");
  if (precision != HGX_PREC_FP64 && precision != HGX_PREC_FP32C)
    return set_error(c, HGX_ERR_INVALID_ARG, "ff_analytic: bad precision");
  Nodes P;
  for (int d = 0; d < 4; ++d) {
    P.n[d] = 0;
    if (d >= kDims[shape]) continue;
    int n = params->n[d];
    if (n < 1 || n > HGX_MAX_NODES || !params->val[d] || !params->wt[d])
      return set_error(c, HGX_ERR_INVALID_ARG, "ff_analytic: dimension %d needs 1..%d nodes with values and weights", d,
                       HGX_MAX_NODES);
    P.n[d] = n;
    for (int i = 0; i < n; ++i) {
      P.val[d][i] = params->val[d][i];
      P.wt[d][i] = params->wt[d][i];
    }
  }
  P.tan_tau = std::tan(params->tau);
  P.sin_eta = std::sin(params->eta);
  P.cos_eta = std::cos(params->eta);
  P.flags = params->flags;
  if (int rc = set_device(c)) return rc;
  QViewA g{c->d_qx, c->d_qy, c->d_qz, c->npix, c->nq};
  Rot R = make_rot(rot);
  Vec3 tv = make_vec(transvec);
  int has_tv = transvec != nullptr;
  unsigned nb = ceil_div(c->nq, 256);
  LaunchTimer lt(c, true);
  switch (shape) {
    case HGX_SHAPE_SPHERE: k_ff_analytic<HGX_SHAPE_SPHERE><<<nb, 256, 0, c->stream>>>(g, R, P, tv, has_tv, ff); break;
    case HGX_SHAPE_CYLINDER: k_ff_analytic<HGX_SHAPE_CYLINDER><<<nb, 256, 0, c->stream>>>(g, R, P, tv, has_tv, ff); break;
    case HGX_SHAPE_BOX: k_ff_analytic<HGX_SHAPE_BOX><<<nb, 256, 0, c->stream>>>(g, R, P, tv, has_tv, ff); break;
    case HGX_SHAPE_CUBE: k_ff_analytic<HGX_SHAPE_CUBE><<<nb, 256, 0, c->stream>>>(g, R, P, tv, has_tv, ff); break;
    case HGX_SHAPE_PRISM3: k_ff_analytic<HGX_SHAPE_PRISM3><<<nb, 256, 0, c->stream>>>(g, R, P, tv, has_tv, ff); break;
    case HGX_SHAPE_PRISM6: k_ff_analytic<HGX_SHAPE_PRISM6><<<nb, 256, 0, c->stream>>>(g, R, P, tv, has_tv, ff); break;
    default: k_ff_analytic<HGX_SHAPE_PYRAMID><<<nb, 256, 0, c->stream>>>(g, R, P, tv, has_tv, ff); break;
  }
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

// AnalyticFormFactor::param_distribution (src/ff/ff_ana.cpp:235-284): node values by the
// reference's accumulating do-while, weights none/uniform/gaussian (unnormalised pdf). Host-only
// arithmetic; returns the node count or a negative error code.
extern "C" int hgx_param_distribution(double pmin, double pmax, int nvalues, int stat, double mean, double sd,
                                      double* vals, double* wts, int cap) {
  if (nvalues < 1 || !vals || !wts) return HGX_ERR_INVALID_ARG;
  double pmx = pmax < pmin ? pmin : pmax;
  int n = 0;
  double node[4096];
  if (nvalues > 1) {
    double step = std::fabs(pmx - pmin) / (nvalues - 1);
    double curr = pmin;
    do {
      if (n >= 4096) return HGX_ERR_INVALID_ARG;
      node[n++] = curr;
      curr += step;
    } while (curr < pmax);
  } else {
    node[n++] = pmin;
  }
  if (stat < 0 || stat > 2) return HGX_ERR_INVALID_ARG;
  double m = mean;
  if (stat == 2 && !std::isfinite(m)) m = (node[0] + node[n - 1]) / 2;
  for (int i = 0; i < n && i < cap; ++i) {
    vals[i] = node[i];
    if (stat == 0) wts[i] = i == 0 ? 1.0 : 0.0;
    else if (stat == 1) wts[i] = 1.0;
    else wts[i] = std::exp(-1.0 * std::pow((node[i] - m), 2) / (2 * std::pow(sd, 2))) / (std::sqrt(2 * kPi) * sd);
  }
  return n;
}
