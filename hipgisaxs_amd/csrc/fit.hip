// fit.hip -- chi^2 of a lattice-of-particles model against an observed image, and its gradient.
//
// Model (one grain): I(p) = s |sum_b fc_b dn2 F(m_b) S(m_b)|^2, m_b = R_x(tilt) rot0 q_b, with
// theta = (radius, height, lattice constant a, tilt, Debye-Waller sigma, scale s); the particle is a
// cylinder (src/ff/ff_ana_cylinder.cpp:34-42) or a sphere (src/ff/ff_ana_sphere.cpp:33-45) and the
// lattice sum is the Dirichlet form of src/sf/sf.cpp:165-245 with vectors a * (ua, ub, uc).
// chi^2 = sum_p mask_p (I_p - O_p)^2 / max(O_p, 1).
//
// The reference fits with TAO's LMVM and a central finite-difference gradient, 2P+1 full
// simulations per gradient (src/analyzer/hipgisaxs_fit_lmvm.cpp:190-224). Here the backward pass is
// written by hand: every thread recomputes its pixel's amplitude and the derivatives of F and S
// with respect to the six parameters (no stored intermediates), forms
// dchi2/dtheta_j = sum_p 2 mask_p (I_p - O_p)/max(O_p,1) dI_p/dtheta_j, and the per-pixel terms
// are reduced per block in LDS and then over blocks in a fixed order (deterministic).
#include <cmath>

#include "hgx_internal.hpp"

namespace hgx {

constexpr double kPiF64 = 3.141592653589793238462643383279502884;
constexpr int kFitWG = 256;
constexpr int kNOut = 7;  // chi2 + 6 gradient components

struct FitArgs {
  int shape;
  int nb;
  double ua[3], ub[3], uc[3];
  double N[3];
  double center[3];
  double dn2r, dn2i;
  double rot0[9];
  int compat_lb2;
  double th[6];  // radius, height, a, tilt, sigma, scale
};

struct QViewF {
  const double* __restrict__ qx;
  const double* __restrict__ qy;
  const double* __restrict__ qz;
  int64_t npix;
};

// D_N(x) and dD/dx for x wrapped into (-pi, pi] (principal branch of the reference's pow)
__device__ inline void dirichlet_d(double x, double N, double* D, double* dD) {
  double w = x - 2.0 * kPiF64 * rint(x * (0.5 / kPiF64));
  if (fabs(w) < 1e-3) {
    *D = N * (1.0 - (N * N - 1.0) * w * w / 24.0);
    *dD = -N * (N * N - 1.0) * w / 12.0;
    return;
  }
  double sh, ch, sn, cn;
  sincos(0.5 * w, &sh, &ch);
  sincos(0.5 * N * w, &sn, &cn);
  *D = sn / sh;
  *dD = (0.5 * N * cn * sh - 0.5 * sn * ch) / (sh * sh);
}

// G(u) = sinc(u) e^{iu} = (e^{2iu} - 1)/(2iu) and dG/du
__device__ inline void sinc_expi(double u, cd* G, cd* dG) {
  if (fabs(u) < 1e-4) {
    // G = 1 + iu - (2/3)u^2 - (i/3)u^3 ; G' = i - (4/3)u - i u^2
    *G = cd{1.0 - (2.0 / 3.0) * u * u, u - u * u * u / 3.0};
    *dG = cd{-(4.0 / 3.0) * u, 1.0 - u * u};
    return;
  }
  double s, c, s2, c2;
  sincos(u, &s, &c);
  sincos(2.0 * u, &s2, &c2);
  *G = cd{s * c / u, s * s / u};  // sin(u)/u (cos u + i sin u)
  // G' = (e^{2iu} - G)/u
  *dG = cd{(c2 - G->r) / u, (s2 - G->i) / u};
}

template <bool GRAD>
__global__ __launch_bounds__(kFitWG) void k_fit(QViewF g, const double* __restrict__ fc, const double* __restrict__ obs,
                                                const double* __restrict__ mask, FitArgs a, double* __restrict__ Imod,
                                                double* __restrict__ part) {
  __shared__ double red[kNOut][kFitWG / 64];
  const int64_t p = (int64_t)blockIdx.x * kFitWG + threadIdx.x;
  double out[kNOut];
#pragma unroll
  for (int k = 0; k < kNOut; ++k) out[k] = 0.0;
  if (p < g.npix) {
    const double R = a.th[0], H = a.th[1], A = a.th[2], t = a.th[3], sig = a.th[4], sc = a.th[5];
    double ct, st;
    sincos(t, &st, &ct);
    const cd dn2 = cd{a.dn2r, a.dn2i};
    cd amp = cd{0.0, 0.0};
    cd dA[5] = {cd{0, 0}, cd{0, 0}, cd{0, 0}, cd{0, 0}, cd{0, 0}};  // d amp / d(R, H, a, t, sigma)
    const double qx = g.qx[p], qy = g.qy[p];
    for (int b = 0; b < a.nb; ++b) {
      const int64_t q = p + b * g.npix;
      const double qz = g.qz[2 * q];
      double m0[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) m0[j] = a.rot0[3 * j] * qx + a.rot0[3 * j + 1] * qy + a.rot0[3 * j + 2] * qz;
      // m = R_x(t) m0 ; dm/dt = (0, -m_z, m_y)
      const double m[3] = {m0[0], ct * m0[1] - st * m0[2], st * m0[1] + ct * m0[2]};
      const double dm[3] = {0.0, -m[2], m[1]};
      // ---- particle form factor and its derivatives
      cd F, dF_R, dF_H = cd{0, 0}, dF_t;
      if (a.shape == HGX_SHAPE_CYLINDER) {
        const double qp = sqrt(m[0] * m[0] + m[1] * m[1]);
        const double x = qp * R;
        double B, dB;  // J1(x)/x and its derivative
        if (x < 1e-3) {
          B = 0.5 - x * x / 16.0;
          dB = -x / 8.0;
        } else {
          double J1 = j1(x);
          B = J1 / x;
          dB = (j0(x) - 2.0 * B) / x;
        }
        cd G, dG;
        sinc_expi(0.5 * H * m[2], &G, &dG);
        const double V = 2.0 * kPiF64 * R * R * H;
        F = (V * B) * G;
        dF_R = V * (2.0 * B / R + qp * dB) * G;
        dF_H = (V * B) * (G * (1.0 / H) + (0.5 * m[2]) * dG);
        const double dqp = qp > 0.0 ? (m[1] * dm[1]) / qp : 0.0;
        dF_t = (V * R * dB * dqp) * G + (V * B * 0.5 * H * dm[2]) * dG;
      } else {
        const double qv = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        const double x = qv * R;
        double gx, dg;  // (sin x - x cos x)/x^3 and its derivative
        if (x < 1e-2) {
          gx = 1.0 / 3.0 - x * x / 30.0 + x * x * x * x / 840.0;
          dg = -x / 15.0 + x * x * x / 210.0;
        } else {
          double s, c;
          sincos(x, &s, &c);
          gx = (s - x * c) / (x * x * x);
          dg = s / (x * x) - 3.0 * gx / x;
        }
        cd e = cexpi_d(cd{m[2] * R, 0.0});
        const double V = 4.0 * kPiF64 * R * R * R;
        F = (qv == 0.0 ? 0.0 : V * gx) * e;  // the reference returns 0 at q = 0
        dF_R = qv == 0.0 ? cd{0, 0} : (4.0 * kPiF64 * (3.0 * R * R * gx + R * R * R * dg * qv)) * e + mul_i(F) * m[2];
        dF_t = mul_i(F) * (R * dm[2]);
      }
      // ---- lattice sum and its derivatives
      const double pa = a.ua[0] * m[0] + a.ua[1] * m[1] + a.ua[2] * m[2];
      const double pb_lin = a.ub[0] * m[0] + a.ub[1] * m[1];
      const double pc = a.uc[0] * m[0] + a.uc[1] * m[1] + a.uc[2] * m[2];
      double xb, dxb_da, dxb_dt;
      if (a.compat_lb2) {  // src/sf/sf.cpp:212 uses lb[2]*lb[2]
        xb = A * pb_lin + A * A * a.ub[2] * a.ub[2];
        dxb_da = pb_lin + 2.0 * A * a.ub[2] * a.ub[2];
        dxb_dt = A * (a.ub[1] * dm[1]);
      } else {
        xb = A * (pb_lin + a.ub[2] * m[2]);
        dxb_da = pb_lin + a.ub[2] * m[2];
        dxb_dt = A * (a.ub[1] * dm[1] + a.ub[2] * dm[2]);
      }
      double Da, dDa, Db, dDb, Dc, dDc;
      dirichlet_d(A * pa, a.N[0], &Da, &dDa);
      dirichlet_d(xb, a.N[1], &Db, &dDb);
      dirichlet_d(A * pc, a.N[2], &Dc, &dDc);
      const double S0 = Da * Db * Dc;
      const double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
      const double W = exp(-0.5 * sig * sig * m2);
      const double cph = a.center[0] * m[0] + a.center[1] * m[1] + a.center[2] * m[2];
      const cd P = cexpi_d(cd{cph, 0.0});
      const cd S = (S0 * W) * P;
      const double dS0_da = dDa * pa * Db * Dc + Da * dDb * dxb_da * Dc + Da * Db * dDc * pc;
      const double dpa_t = a.ua[1] * dm[1] + a.ua[2] * dm[2], dpc_t = a.uc[1] * dm[1] + a.uc[2] * dm[2];
      const double dS0_dt = dDa * A * dpa_t * Db * Dc + Da * dDb * dxb_dt * Dc + Da * Db * dDc * A * dpc_t;
      const double dcph_t = a.center[1] * dm[1] + a.center[2] * dm[2];
      const cd dS_da = (dS0_da * W) * P;
      const cd dS_dt = (dS0_dt * W) * P + mul_i(S) * dcph_t;
      const cd dS_ds = (-sig * m2) * S;
      // ---- amplitude
      const cd f = fc ? dn2 * cd{fc[2 * q], fc[2 * q + 1]} : dn2;
      amp = amp + f * (F * S);
      if (GRAD) {
        dA[0] = dA[0] + f * (dF_R * S);
        dA[1] = dA[1] + f * (dF_H * S);
        dA[2] = dA[2] + f * (F * dS_da);
        dA[3] = dA[3] + f * (dF_t * S + F * dS_dt);
        dA[4] = dA[4] + f * (F * dS_ds);
      }
    }
    const double A2 = norm2(amp);
    const double I = sc * A2;
    if (Imod) Imod[p] = I;
    const double o = obs ? obs[p] : 0.0;
    const double mk = mask ? mask[p] : 1.0;
    const double wgt = mk / fmax(o, 1.0);
    out[0] = wgt * (I - o) * (I - o);
    if (GRAD) {
      const double gI = 2.0 * wgt * (I - o);
#pragma unroll
      for (int j = 0; j < 5; ++j) out[1 + j] = gI * 2.0 * sc * (amp.r * dA[j].r + amp.i * dA[j].i);
      out[6] = gI * A2;
    }
  }
  // block reduction: wave shuffles, then the 4 wave sums through LDS, in a fixed order
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kNOut; ++k) {
    double v = out[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[k][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < kNOut) {
    double s = 0.0;
    for (int w = 0; w < kFitWG / 64; ++w) s += red[threadIdx.x][w];
    part[(int64_t)blockIdx.x * kNOut + threadIdx.x] = s;
  }
}

// Sums the per-block partials: one workgroup of 256 threads; thread j adds blocks j, j+256, ... for
// every output, then a fixed-shape LDS tree combines the 256 sums (same order on every run).
__global__ __launch_bounds__(256) void k_fit_finish(const double* __restrict__ part, int nblocks,
                                                    double* __restrict__ out) {
  __shared__ double red[kNOut][256];
  const int j = threadIdx.x;
  double s[kNOut];
#pragma unroll
  for (int k = 0; k < kNOut; ++k) s[k] = 0.0;
  for (int b = j; b < nblocks; b += 256)
#pragma unroll
    for (int k = 0; k < kNOut; ++k) s[k] += part[(int64_t)b * kNOut + k];
#pragma unroll
  for (int k = 0; k < kNOut; ++k) red[k][j] = s[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (j < w)
#pragma unroll
      for (int k = 0; k < kNOut; ++k) red[k][j] += red[k][j + w];
    __syncthreads();
  }
  if (j < kNOut) out[j] = red[j][0];
}

static std::vector<FitPlan>& plans(Context* c) { return c->fit_plans; }

static int run_fit(Context* c, int plan, const double theta[6], int precision, bool grad, double* chi2,
                   double* gradient, double* Imod) {
  auto& P = plans(c);
  if (plan < 0 || plan >= (int)P.size() || !P[plan].live) return set_error(c, HGX_ERR_BAD_HANDLE, "fit: unknown plan %d", plan);
  if (!theta || !chi2) return set_error(c, HGX_ERR_INVALID_ARG, "fit: null theta or chi2");
  if (precision != HGX_PREC_FP64) return set_error(c, HGX_ERR_UNSUPPORTED, "fit: FP64 only");
  if (!c->qz_real) return set_error(c, HGX_ERR_UNSUPPORTED, "fit: needs real q (non-absorbing layer)");
  const FitPlan& fp = P[plan];
  const hgx_fit_model& m = fp.model;
  if (!(theta[0] > 0) || (m.shape == HGX_SHAPE_CYLINDER && !(theta[1] > 0)) || !(theta[2] > 0))
    return set_error(c, HGX_ERR_INVALID_ARG, "fit: radius, height and lattice constant must be positive");
  FitArgs a;
  a.shape = m.shape;
  a.nb = c->nbranch;
  for (int k = 0; k < 3; ++k) {
    a.ua[k] = m.ua[k];
    a.ub[k] = m.ub[k];
    a.uc[k] = m.uc[k];
    a.N[k] = m.repet[k] < 1 ? 1 : m.repet[k];
    a.center[k] = m.center[k];
  }
  for (int k = 0; k < 9; ++k) a.rot0[k] = m.rot0[k];
  a.dn2r = m.dn2_re;
  a.dn2i = m.dn2_im;
  a.compat_lb2 = (m.sf_flags & HGX_SF_COMPAT_LB2) != 0;
  for (int k = 0; k < 6; ++k) a.th[k] = theta[k];
  unsigned nblk = ceil_div(c->npix, kFitWG);
  if (int rc = ensure_work(c, ((size_t)nblk * kNOut + kNOut) * sizeof(double))) return rc;
  double* part = reinterpret_cast<double*>(c->d_work);
  double* res = part + (size_t)nblk * kNOut;
  QViewF g{c->d_qx, c->d_qy, c->d_qz, c->npix};
  {
    LaunchTimer lt(c, true);
    if (grad)
      k_fit<true><<<nblk, kFitWG, 0, c->stream>>>(g, fp.fc, fp.obs, fp.mask, a, Imod, part);
    else
      k_fit<false><<<nblk, kFitWG, 0, c->stream>>>(g, fp.fc, fp.obs, fp.mask, a, Imod, part);
  }
  HGX_CHECK_LAUNCH(c);
  {
    LaunchTimer lt(c, false);
    k_fit_finish<<<1, 256, 0, c->stream>>>(part, (int)nblk, res);
  }
  HGX_CHECK_LAUNCH(c);
  double host[kNOut];
  HGX_CHECK_HIP(c, hipMemcpyAsync(host, res, sizeof(host), hipMemcpyDeviceToHost, c->stream));
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  *chi2 = host[0];
  if (grad && gradient) {
    // kernel order: R, H, a, tilt, sigma, scale
    for (int k = 0; k < 6; ++k) gradient[k] = host[1 + k];
    if (m.shape == HGX_SHAPE_SPHERE) gradient[1] = 0.0;
  }
  return HGX_OK;
}

}  // namespace hgx

using namespace hgx;

extern "C" int hgx_fit_plan_create(hgx_ctx* h, const hgx_fit_model* model, const double* fc, const double* I_obs,
                                   const double* mask, int* plan) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "fit_plan_create: no q-grid");
  if (!model || !plan) return set_error(c, HGX_ERR_INVALID_ARG, "fit_plan_create: null argument");
  if (model->shape != HGX_SHAPE_CYLINDER && model->shape != HGX_SHAPE_SPHERE)
    return set_error(c, HGX_ERR_UNSUPPORTED, "fit_plan_create: shape must be cylinder or sphere");
  auto& P = plans(c);
  FitPlan fp;
  fp.live = true;
  fp.model = *model;
  fp.fc = fc;
  fp.obs = I_obs;
  fp.mask = mask;
  for (size_t i = 0; i < P.size(); ++i)
    if (!P[i].live) {
      P[i] = fp;
      *plan = (int)i;
      return HGX_OK;
    }
  P.push_back(fp);
  *plan = (int)P.size() - 1;
  return HGX_OK;
}

extern "C" int hgx_fit_plan_release(hgx_ctx* h, int plan) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  auto& P = plans(c);
  if (plan < 0 || plan >= (int)P.size() || !P[plan].live) return set_error(c, HGX_ERR_BAD_HANDLE, "fit: unknown plan");
  P[plan].live = false;
  return HGX_OK;
}

extern "C" int hgx_chi2_fwd(hgx_ctx* h, int plan, const double theta[6], int precision, double* chi2, double* I_model) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (int rc = set_device(c)) return rc;
  return run_fit(c, plan, theta, precision, false, chi2, nullptr, I_model);
}

extern "C" int hgx_chi2_bwd(hgx_ctx* h, int plan, const double theta[6], int precision, double* chi2, double grad[6]) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!grad) return set_error(c, HGX_ERR_INVALID_ARG, "chi2_bwd: null gradient");
  if (int rc = set_device(c)) return rc;
  return run_fit(c, plan, theta, precision, true, chi2, grad, nullptr);
}
