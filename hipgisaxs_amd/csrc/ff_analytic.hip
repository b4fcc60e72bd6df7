// ff_analytic.hip -- analytic form factors, one thread per q-point, node loops in registers: plain
// FP64 (analytic_shapes.hpp) or FP32C on real q (analytic32.hpp: FP64-reduced phases, FP32 sines and
// node sums, the FP64 formula where the prism and pyramid closed forms cancel).
// Formulas are the CPU ones of the reference (src/ff/ff_ana_<shape>.cpp), which the reference's
// GPU kernels do not all follow (src/ff/gpu/ff_ana_prism_gpu.cu:39-57, ff_ana_prism6_gpu.cu:47).
#include <cmath>

#include "analytic32.hpp"
#include "analytic_shapes.hpp"

namespace hgx {

struct QViewA {
  const double* __restrict__ qx;
  const double* __restrict__ qy;
  const double* __restrict__ qz;
  int64_t npix, nq;
};

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

template <int SHAPE>
__global__ __launch_bounds__(256) void k_ff_analytic32(QViewA g, Rot R, Nodes P, Vec3 tv, int has_tv,
                                                       double* __restrict__ out) {
  int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= g.nq) return;
  int64_t p = q % g.npix;
  const double qx = g.qx[p], qy = g.qy[p], qz = g.qz[2 * q];
  double m[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) m[j] = R.m[3 * j] * qx + R.m[3 * j + 1] * qy + R.m[3 * j + 2] * qz;
  cd F = ff_fp32c<SHAPE>(m, P);
  if (has_tv) {
    const cf e = expi_turn(turns_of(m[0] * tv.v[0] + m[1] * tv.v[1] + m[2] * tv.v[2]));
    F = F * cd{(double)e.r, (double)e.i};
  }
  out[2 * q] = F.r;
  out[2 * q + 1] = F.i;
}

}  // namespace hgx

using namespace hgx;

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
    if (d >= analytic_dims(shape, params->flags)) continue;
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
  if (!analytic_heights_ok(shape, params->flags, params->n, params->val))
    return set_error(c, HGX_ERR_INVALID_ARG, "ff_analytic: truncated sphere heights must lie in [0, 2 R]");
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
  // spheres without a cap run the lean instantiation (no truncated-sphere code)
  const bool plain_sphere = shape == HGX_SHAPE_SPHERE && !((params->flags & HGX_SHAPE_TRUNCATED_SPHERE) && P.n[1] > 0);
  if (precision == HGX_PREC_FP32C) {
    if (!c->qz_real)
      return set_error(c, HGX_ERR_UNSUPPORTED, "ff_analytic: FP32C needs real q (a particle in a non-absorbing layer)");
    LaunchTimer lt(c, true);
#define HGX_A32(S) k_ff_analytic32<S><<<nb, 256, 0, c->stream>>>(g, R, P, tv, has_tv, ff)
    switch (shape) {
      case HGX_SHAPE_SPHERE:
        if (plain_sphere) HGX_A32(kShapePlainSphere);
        else HGX_A32(HGX_SHAPE_SPHERE);
        break;
      case HGX_SHAPE_CYLINDER: HGX_A32(HGX_SHAPE_CYLINDER); break;
      case HGX_SHAPE_BOX: HGX_A32(HGX_SHAPE_BOX); break;
      case HGX_SHAPE_CUBE: HGX_A32(HGX_SHAPE_CUBE); break;
      case HGX_SHAPE_PRISM3: HGX_A32(HGX_SHAPE_PRISM3); break;
      case HGX_SHAPE_PRISM6: HGX_A32(HGX_SHAPE_PRISM6); break;
      default: HGX_A32(HGX_SHAPE_PYRAMID); break;
    }
#undef HGX_A32
    HGX_CHECK_LAUNCH(c);
    return HGX_OK;
  }
  LaunchTimer lt(c, true);
  switch (shape) {
    case HGX_SHAPE_SPHERE:
      if (plain_sphere) k_ff_analytic<kShapePlainSphere><<<nb, 256, 0, c->stream>>>(g, R, P, tv, has_tv, ff);
      else k_ff_analytic<HGX_SHAPE_SPHERE><<<nb, 256, 0, c->stream>>>(g, R, P, tv, has_tv, ff);
      break;
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

// AnalyticFormFactor::param_distribution (src/ff/ff_ana.cpp:235-284). Default: exactly nvalues
// nodes on [pmin, pmax] with weights normalised to sum 1. HGX_DIST_REFERENCE: the reference's
// accumulating do-while (nvalues +- 1 nodes) and unnormalised pdf weights. Host-only arithmetic;
// returns the node count or a negative error code.
extern "C" int hgx_param_distribution(double pmin, double pmax, int nvalues, int stat, double mean, double sd,
                                      int flags, double* vals, double* wts, int cap) {
  if (nvalues < 1 || !vals || !wts) return HGX_ERR_INVALID_ARG;
  if (flags & ~HGX_DIST_REFERENCE) return HGX_ERR_INVALID_ARG;
  if (stat < 0 || stat > 2) return HGX_ERR_INVALID_ARG;
  const bool reference = (flags & HGX_DIST_REFERENCE) != 0;
  double pmx = pmax < pmin ? pmin : pmax;
  int n = 0;
  double node[4096];
  if (nvalues > 1 && reference) {
    double step = std::fabs(pmx - pmin) / (nvalues - 1);
    double curr = pmin;
    do {
      if (n >= 4096) return HGX_ERR_INVALID_ARG;
      node[n++] = curr;
      curr += step;
    } while (curr < pmax);
  } else if (nvalues > 1) {
    if (nvalues > 4096) return HGX_ERR_INVALID_ARG;
    for (int k = 0; k < nvalues; ++k) node[n++] = pmin + (pmx - pmin) * k / (nvalues - 1);
  } else {
    node[n++] = pmin;
  }
  double m = mean;
  if (stat == 2 && !std::isfinite(m)) m = (node[0] + node[n - 1]) / 2;
  double wsum = 0.0;
  if (stat == 2 && !reference) {
    for (int i = 0; i < n; ++i)
      wsum += std::exp(-1.0 * std::pow((node[i] - m), 2) / (2 * std::pow(sd, 2))) / (std::sqrt(2 * kPi) * sd);
    if (!(wsum > 0)) return HGX_ERR_INVALID_ARG;
  }
  for (int i = 0; i < n && i < cap; ++i) {
    vals[i] = node[i];
    if (stat == 0) {
      wts[i] = i == 0 ? 1.0 : 0.0;
    } else if (stat == 1) {
      wts[i] = reference ? 1.0 : 1.0 / n;
    } else {
      double pdf = std::exp(-1.0 * std::pow((node[i] - m), 2) / (2 * std::pow(sd, 2))) / (std::sqrt(2 * kPi) * sd);
      wts[i] = reference ? pdf : pdf / wsum;
    }
  }
  return n;
}
