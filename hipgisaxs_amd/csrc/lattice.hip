// lattice.hip -- fused intensity of a grain ensemble of lattices of analytic particles (DWBA).
//
// For every pixel p (one thread), grain orientation g and DWBA branch b:
//     I[p] (+)= sum_g | w_g dn2 sum_b fc_b F(R_g q_b) S(R_g q_b) |^2
// which is the grain loop, the unit-cell element sum (one element), the structure factor and the
// DWBA combination of HipGISAXS::run_gisaxs (src/sim/hipgisaxs_main.cpp:925-1100) without the
// intermediate complex [4 npix] arrays the reference writes and re-reads per grain.
//
// FP64 plain: the same FP64 device functions as the per-shape / SF kernels (analytic_shapes.hpp,
// lattice_sf.hpp), so it matches them and the oracle to round-off.
// FP32C: q rotated and lattice phases reduced in FP64; the node loops and Dirichlet kernels in FP32
// with turn-based hardware sin/cos and Taylor sines at small arguments; amplitudes summed and
// squared in FP64. Real q only (a particle in a non-absorbing layer).
#include <cmath>

#include "analytic_shapes.hpp"
#include "fp32_math.hpp"
#include "lattice_sf.hpp"

namespace hgx {

constexpr int kMaxGrains = 16;

struct LatArgs {
  int shape;
  int nb;  // DWBA branches (1 or 4)
  int n0, n1;
  double v0[HGX_MAX_NODES], w0[HGX_MAX_NODES];  // sphere: radius; cylinder: radius
  double v1[HGX_MAX_NODES], w1[HGX_MAX_NODES];  // cylinder: height
  SFArgs sf;
  double dn2r, dn2i;
  int ng;
  double rot[kMaxGrains][9];
  double wg[kMaxGrains];
};

struct QViewL {
  const double* __restrict__ qx;
  const double* __restrict__ qy;
  const double* __restrict__ qz;
  int64_t npix;
};

__device__ inline void rot_g(const LatArgs& a, int g, double qx, double qy, cd qz, cd mq[3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    mq[j].r = a.rot[g][3 * j] * qx + a.rot[g][3 * j + 1] * qy + a.rot[g][3 * j + 2] * qz.r;
    mq[j].i = a.rot[g][3 * j + 2] * qz.i;
  }
}

// ----- FP64 particle form factors (the formulas of analytic_shapes.hpp on the packed nodes) -----

__device__ cd sphere64(const cd mq[3], const LatArgs& a) {
  cd qval = csqrt_d(mq[0] * mq[0] + mq[1] * mq[1] + mq[2] * mq[2]);
  cd acc = cd{0.0, 0.0};
  if (cabs_d(qval) < 2.2250738585072014e-308) return acc;  // src/ff/ff_ana_sphere.cpp:35-37
  for (int i = 0; i < a.n0; ++i) {
    double R = a.v0[i];
    cd qR = qval * R;
    cd qR3 = qR * qR * qR;
    cd c0 = (csin_d(qR) - qR * ccos_d(qR)) * ((1.0 / norm2(qR3)) * conj(qR3));
    acc = acc + (a.w0[i] * (4 * kPi * R * R * R)) * (c0 * cexpi_d(mq[2] * R));
  }
  return acc;
}

// cylinder (src/ff/ff_ana_cylinder.cpp:34-42) with the node double sum factorised:
// sum_{r,h} w_r w_h V(r,h) B(r) G(h) = [sum_r w_r 2 pi r^2 B(r)] [sum_h w_h h sinc(u_h) e^{i u_h}]
__device__ cd cylinder64(const cd mq[3], const LatArgs& a) {
  cd qpar = csqrt_d(mq[0] * mq[0] + mq[1] * mq[1]);
  cd br = cd{0.0, 0.0};
  for (int i = 0; i < a.n0; ++i) {
    double R = a.v0[i];
    cd t1 = qpar * R;
    double nt = norm2(t1);
    if (nt > 1.0e-13) br = br + (a.w0[i] * 2.0 * kPi * R * R * j1(t1.r) / nt) * conj(t1);
  }
  cd hz = cd{0.0, 0.0};
  for (int k = 0; k < a.n1; ++k) {
    double H = a.v1[k];
    cd u = (0.5 * H) * mq[2];
    hz = hz + (a.w1[k] * H) * (sinc_free(u) * cexpi_d(u));
  }
  return br * hz;
}

__global__ __launch_bounds__(256) void k_lattice64(QViewL g, const double* __restrict__ fc, LatArgs a, int accumulate,
                                                   double* __restrict__ I) {
  int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const double qx = g.qx[p], qy = g.qy[p];
  // the pixel's branch data, read once for all grains
  // (fixed-bound unrolled loops keep the arrays in registers)
  cd qzb[4], fcb[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if (b >= a.nb) break;
    int64_t q = p + b * g.npix;
    qzb[b] = cd{g.qz[2 * q], g.qz[2 * q + 1]};
    fcb[b] = fc ? cd{fc[2 * q], fc[2 * q + 1]} : cd{1.0, 0.0};
  }
  double acc = 0.0;
  for (int o = 0; o < a.ng; ++o) {
    cd amp = cd{0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (b >= a.nb) break;
      cd mq[3];
      rot_g(a, o, qx, qy, qzb[b], mq);
      cd F = a.shape == HGX_SHAPE_SPHERE ? sphere64(mq, a) : cylinder64(mq, a);
      cd S = sf_value(mq, a.sf);
      amp = amp + fcb[b] * F * S;
    }
    amp = a.wg[o] * (cd{a.dn2r, a.dn2i} * amp);
    acc += norm2(amp);
  }
  I[p] = accumulate ? I[p] + acc : acc;
}

// ----- FP32C -----------------------------------------------------------------------------------

// sum_k w_k 4 pi R_k^3 (sin x - x cos x)/x^3 e^{i qz R_k}, x = |q| R_k, with the node-independent
// 1/|q|^3 taken out of the loop; a 4-term series replaces sin x - x cos x below x = 0.6.
__device__ cf sphere32(float qv, float mz, const LatArgs& a) {
  if (qv == 0.0f) return cf{0.0f, 0.0f};
  const float qt = qv * (1.0f / kTwoPiF), mzt = mz * (1.0f / kTwoPiF);
  cf acc = cf{0.0f, 0.0f};
  for (int i = 0; i < a.n0; ++i) {
    float R = (float)a.v0[i];
    float x = qv * R;
    float t;
    if (x < 0.6f) {
      float x2 = x * x;
      t = x2 * x * ((1.0f / 3.0f) - x2 * ((1.0f / 30.0f) - x2 * ((1.0f / 840.0f) - x2 * (1.0f / 45360.0f))));
    } else {
      float tr = qt * R;
      t = sin_turn(tr) - x * cos_turn(tr);
    }
    float c = (float)(a.w0[i] * 4.0 * kPi) * t;
    cf e = expi_turn(mzt * R);
    acc = acc + c * e;
  }
  float iq3 = 1.0f / (qv * qv * qv);
  return iq3 * acc;
}

__device__ cf cylinder32(float qp, float mz, const LatArgs& a) {
  float br = 0.0f;
  for (int i = 0; i < a.n0; ++i) {
    float R = (float)a.v0[i];
    if (qp * R * qp * R > 1.0e-13f) br += (float)(a.w0[i] * 2.0 * kPi) * R * R * j1_over_x(qp * R);
  }
  cf hz = cf{0.0f, 0.0f};
  for (int k = 0; k < a.n1; ++k) {
    float H = (float)a.v1[k];
    float u = 0.5f * H * mz;
    float s;
    if (fabsf(u) < 0.1f) {
      float u2 = u * u;
      s = 1.0f - u2 * (1.0f / 6.0f) * (1.0f - u2 * (1.0f / 20.0f));
    } else {
      s = sin_turn(u * (1.0f / kTwoPiF)) / u;
    }
    hz = hz + ((float)a.w1[k] * H * s) * expi_turn(u * (1.0f / kTwoPiF));
  }
  return br * hz;
}

// lattice phase l.q reduced to (-1/2, 1/2] turns in FP64
__device__ inline float wrap_turns(double x) {
  double u = x * (1.0 / (2.0 * kPiD));
  return (float)(u - rint(u));
}

__device__ cf sf32(const double m[3], const LatArgs& a) {
  const SFArgs& s = a.sf;
  double xa = m[0] * s.la[0] + m[1] * s.la[1] + m[2] * s.la[2];
  double xb = m[0] * s.lb[0] + m[1] * s.lb[1] + (s.compat_lb2 ? s.lb[2] * s.lb[2] : m[2] * s.lb[2]);
  double xc = m[0] * s.lc[0] + m[1] * s.lc[1] + m[2] * s.lc[2];
  float d = dirichlet_turn(wrap_turns(xa), (float)s.N[0]) * dirichlet_turn(wrap_turns(xb), (float)s.N[1]) *
            dirichlet_turn(wrap_turns(xc), (float)s.N[2]);
  cf ph = expi_turn(wrap_turns(m[0] * s.center[0] + m[1] * s.center[1] + m[2] * s.center[2]));
  if (s.dw != 0.0) d *= __expf((float)(-s.dw * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2])));
  return d * ph;
}

__global__ __launch_bounds__(256) void k_lattice32(QViewL g, const double* __restrict__ fc, LatArgs a, int accumulate,
                                                   double* __restrict__ I) {
  int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const double qx = g.qx[p], qy = g.qy[p];
  // the pixel's branch data, read once for all grains
  // (fixed-bound unrolled loops keep the arrays in registers)
  double qzb[4];
  cd fcb[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if (b >= a.nb) break;
    int64_t q = p + b * g.npix;
    qzb[b] = g.qz[2 * q];
    fcb[b] = fc ? cd{fc[2 * q], fc[2 * q + 1]} : cd{1.0, 0.0};
  }
  double acc = 0.0;
  for (int o = 0; o < a.ng; ++o) {
    cd amp = cd{0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (b >= a.nb) break;
      const double qz = qzb[b];
      double m[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) m[j] = a.rot[o][3 * j] * qx + a.rot[o][3 * j + 1] * qy + a.rot[o][3 * j + 2] * qz;
      cf F;
      if (a.shape == HGX_SHAPE_SPHERE)
        F = sphere32((float)sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]), (float)m[2], a);
      else
        F = cylinder32((float)sqrt(m[0] * m[0] + m[1] * m[1]), (float)m[2], a);
      cf FS = F * sf32(m, a);
      amp = amp + fcb[b] * cd{(double)FS.r, (double)FS.i};
    }
    amp = a.wg[o] * (cd{a.dn2r, a.dn2i} * amp);
    acc += norm2(amp);
  }
  I[p] = accumulate ? I[p] + acc : acc;
}

// Packs the model into kernel arguments; returns an error code.
int pack_lattice(Context* c, const hgx_lattice_model* m, int grain0, int ngr, LatArgs* a) {
  if (m->shape != HGX_SHAPE_SPHERE && m->shape != HGX_SHAPE_CYLINDER)
    return set_error(c, HGX_ERR_UNSUPPORTED, "lattice: shape %d not supported (sphere, cylinder)", m->shape);
  int ndim = m->shape == HGX_SHAPE_SPHERE ? 1 : 2;
  a->shape = m->shape;
  a->nb = c->nbranch;
  a->n0 = a->n1 = 0;
  for (int d = 0; d < ndim; ++d) {
    int n = m->params.n[d];
    if (n < 1 || n > HGX_MAX_NODES || !m->params.val[d] || !m->params.wt[d])
      return set_error(c, HGX_ERR_INVALID_ARG, "lattice: dimension %d needs 1..%d nodes", d, HGX_MAX_NODES);
    double* v = d == 0 ? a->v0 : a->v1;
    double* w = d == 0 ? a->w0 : a->w1;
    for (int i = 0; i < n; ++i) {
      v[i] = m->params.val[d][i];
      w[i] = m->params.wt[d][i];
    }
    (d == 0 ? a->n0 : a->n1) = n;
  }
  for (int k = 0; k < 3; ++k) {
    a->sf.la[k] = m->la[k];
    a->sf.lb[k] = m->lb[k];
    a->sf.lc[k] = m->lc[k];
    a->sf.N[k] = m->repet[k] < 1 ? 1 : m->repet[k];
    a->sf.center[k] = m->center[k];
  }
  a->sf.dw = 0.5 * m->dw_sigma * m->dw_sigma;
  a->sf.compat_lb2 = (m->sf_flags & HGX_SF_COMPAT_LB2) != 0;
  a->dn2r = m->dn2_re;
  a->dn2i = m->dn2_im;
  a->ng = ngr;
  for (int o = 0; o < ngr; ++o) {
    for (int k = 0; k < 9; ++k) a->rot[o][k] = m->rot[9 * (grain0 + o) + k];
    a->wg[o] = m->weight ? m->weight[grain0 + o] : 1.0;
  }
  return HGX_OK;
}

}  // namespace hgx

using namespace hgx;

extern "C" int hgx_lattice_intensity(hgx_ctx* h, const hgx_lattice_model* m, int precision, const double* fc,
                                     int accumulate, double* I) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "lattice_intensity: no q-grid");
  if (!m || !I || !m->rot || m->ngrains < 1)
    return set_error(c, HGX_ERR_INVALID_ARG, "lattice_intensity: need a model with >= 1 grain and an output");
  if (precision != HGX_PREC_FP64 && precision != HGX_PREC_FP32C)
    return set_error(c, HGX_ERR_INVALID_ARG, "lattice_intensity: bad precision");
  if (precision == HGX_PREC_FP32C && !c->qz_real)
    return set_error(c, HGX_ERR_UNSUPPORTED, "lattice_intensity: FP32C needs real q (non-absorbing layer)");
  if (int rc = set_device(c)) return rc;
  QViewL g{c->d_qx, c->d_qy, c->d_qz, c->npix};
  for (int g0 = 0; g0 < m->ngrains; g0 += kMaxGrains) {
    int ngr = std::min(kMaxGrains, m->ngrains - g0);
    LatArgs a;
    if (int rc = pack_lattice(c, m, g0, ngr, &a)) return rc;
    int acc = (g0 > 0) ? 1 : accumulate;
    LaunchTimer lt(c, true);
    if (precision == HGX_PREC_FP64)
      k_lattice64<<<ceil_div(c->npix, 256), 256, 0, c->stream>>>(g, fc, a, acc, I);
    else
      k_lattice32<<<ceil_div(c->npix, 256), 256, 0, c->stream>>>(g, fc, a, acc, I);
    HGX_CHECK_LAUNCH(c);
  }
  return HGX_OK;
}
