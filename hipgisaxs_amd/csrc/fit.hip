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
//
// Other objectives (hgx_fit_plan_metric): the reference's distance functors
// (include/analyzer/distance_functions.hpp) -- the masked squared difference
// (AbsoluteDifferenceSquare, :125-143) and the two square-root normalised forms
// (SqrtUnitVectorNormL2DistanceSquare :423-456, SqrtCNormL2DistanceSquare :460-499). The normalised
// forms depend on image-wide sums, so they take a statistics pass (six masked sums) before the
// gradient pass, whose per-pixel seed dD/dI_p is formed from those sums (derivation in the kernel).
#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

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
  int metric;    // HGX_METRIC_*
  double k[3];   // constants of the normalised metrics' per-pixel seed (from the statistics pass)
};

// kernel modes
constexpr int kModeValue = 0;  // objective terms (chi^2, L2) in slot 0
constexpr int kModeGrad = 1;   // objective terms in slot 0 and gradient terms in slots 1..6
constexpr int kModeStats = 2;  // six masked sums of the normalised metrics in slots 0..5

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

template <int MODE>
__global__ __launch_bounds__(kFitWG) void k_fit(QViewF g, const double* __restrict__ fc, const double* __restrict__ obs,
                                                const double* __restrict__ mask, FitArgs a, double* __restrict__ Imod,
                                                double* __restrict__ part) {
  __shared__ double red[kNOut][kFitWG / 64];
  const int64_t p = (int64_t)blockIdx.x * kFitWG + threadIdx.x;
  constexpr bool GRAD = MODE == kModeGrad;
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
    // square roots of model and data for the normalised metrics (a negative value contributes 0, as
    // the reference's NaN guards make it)
    const double sI = I > 0.0 ? sqrt(I) : 0.0, sO = o > 0.0 ? sqrt(o) : 0.0;
    if (MODE == kModeStats) {
      // A1 = sum m I, A2 = sum m sqrt(IO), A3 = sum m O, B1 = sum m^2 I, B2 = sum m^2 sqrt(IO),
      // B3 = sum m^2 O (norm_l2 and vec_dot weight by the mask, the squared distance by its square)
      out[0] = mk * sI * sI;
      out[1] = mk * sI * sO;
      out[2] = mk * sO * sO;
      out[3] = mk * mk * sI * sI;
      out[4] = mk * mk * sI * sO;
      out[5] = mk * mk * sO * sO;
    } else {
      // dD/dI_p and the pixel's share of D (metric_from_sums documents the normalised forms)
      const double gI = metric_seed(a.metric, I, o, mk, a.k, &out[0]);
      if (GRAD) {
#pragma unroll
        for (int j = 0; j < 5; ++j) out[1 + j] = gI * 2.0 * sc * (amp.r * dA[j].r + amp.i * dA[j].i);
        out[6] = gI * A2;
      }
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

static int check_plan(Context* c, int plan) {
  auto& P = plans(c);
  if (plan < 0 || plan >= (int)P.size() || !P[plan].live) return set_error(c, HGX_ERR_BAD_HANDLE, "fit: unknown plan %d", plan);
  if (!c->qz_real) return set_error(c, HGX_ERR_UNSUPPORTED, "fit: needs real q (non-absorbing layer)");
  return HGX_OK;
}

static int pack_fit(Context* c, const FitPlan& fp, const double theta[6], FitArgs* a) {
  const hgx_fit_model& m = fp.model;
  if (!(theta[0] > 0) || (m.shape == HGX_SHAPE_CYLINDER && !(theta[1] > 0)) || !(theta[2] > 0))
    return set_error(c, HGX_ERR_INVALID_ARG, "fit: radius, height and lattice constant must be positive");
  a->shape = m.shape;
  a->nb = c->nbranch;
  for (int k = 0; k < 3; ++k) {
    a->ua[k] = m.ua[k];
    a->ub[k] = m.ub[k];
    a->uc[k] = m.uc[k];
    a->N[k] = m.repet[k] < 1 ? 1 : m.repet[k];
    a->center[k] = m.center[k];
  }
  for (int k = 0; k < 9; ++k) a->rot0[k] = m.rot0[k];
  a->dn2r = m.dn2_re;
  a->dn2i = m.dn2_im;
  a->compat_lb2 = (m.sf_flags & HGX_SF_COMPAT_LB2) != 0;
  for (int k = 0; k < 6; ++k) a->th[k] = theta[k];
  a->metric = fp.metric;
  a->k[0] = a->k[1] = a->k[2] = 0.0;
  return HGX_OK;
}

// One pass of k_fit in `mode` plus the fixed-order block reduction into res[kNOut] (device).
static int launch_fit(Context* c, const FitPlan& fp, const FitArgs& a, int mode, double* Imod, double* part,
                      double* res) {
  unsigned nblk = ceil_div(c->npix, kFitWG);
  QViewF g{c->d_qx, c->d_qy, c->d_qz, c->npix};
  {
    LaunchTimer lt(c, true);
    if (mode == kModeGrad)
      k_fit<kModeGrad><<<nblk, kFitWG, 0, c->stream>>>(g, fp.fc, fp.obs, fp.mask, a, Imod, part);
    else if (mode == kModeStats)
      k_fit<kModeStats><<<nblk, kFitWG, 0, c->stream>>>(g, fp.fc, fp.obs, fp.mask, a, Imod, part);
    else
      k_fit<kModeValue><<<nblk, kFitWG, 0, c->stream>>>(g, fp.fc, fp.obs, fp.mask, a, Imod, part);
  }
  HGX_CHECK_LAUNCH(c);
  {
    LaunchTimer lt(c, false);
    k_fit_finish<<<1, 256, 0, c->stream>>>(part, (int)nblk, res);
  }
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

static bool normalised(int metric) {
  return metric == HGX_METRIC_SQRT_UNIT_L2 || metric == HGX_METRIC_SQRT_C_L2;
}

static int run_fit(Context* c, int plan, const double theta[6], int precision, bool grad, double* chi2,
                   double* gradient, double* Imod) {
  if (int rc = check_plan(c, plan)) return rc;
  if (!theta || !chi2) return set_error(c, HGX_ERR_INVALID_ARG, "fit: null theta or chi2");
  if (precision != HGX_PREC_FP64) return set_error(c, HGX_ERR_UNSUPPORTED, "fit: FP64 only");
  const FitPlan& fp = plans(c)[plan];
  FitArgs a;
  if (int rc = pack_fit(c, fp, theta, &a)) return rc;
  unsigned nblk = ceil_div(c->npix, kFitWG);
  if (int rc = ensure_work(c, ((size_t)nblk * kNOut + 2 * kNOut) * sizeof(double))) return rc;
  double* part = reinterpret_cast<double*>(c->d_work);
  double* res = part + (size_t)nblk * kNOut;
  double host[2 * kNOut];
  const bool norm = normalised(fp.metric);
  if (norm) {
    if (int rc = launch_fit(c, fp, a, kModeStats, Imod, part, res)) return rc;
    HGX_CHECK_HIP(c, hipMemcpyAsync(host, res, kNOut * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    *chi2 = metric_from_sums(fp.metric, host, a.k);
    if (!grad) return HGX_OK;
    if (int rc = launch_fit(c, fp, a, kModeGrad, nullptr, part, res + kNOut)) return rc;
  } else {
    if (int rc = launch_fit(c, fp, a, grad ? kModeGrad : kModeValue, Imod, part, res + kNOut)) return rc;
  }
  HGX_CHECK_HIP(c, hipMemcpyAsync(host + kNOut, res + kNOut, kNOut * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  if (!norm) *chi2 = host[kNOut];
  if (grad && gradient) {
    // kernel order: R, H, a, tilt, sigma, scale
    for (int k = 0; k < 6; ++k) gradient[k] = host[kNOut + 1 + k];
    if (fp.model.shape == HGX_SHAPE_SPHERE) gradient[1] = 0.0;
  }
  return HGX_OK;
}

// Objective of n parameter vectors queued back to back, one synchronisation at the end.
static int run_fit_many(Context* c, int plan, int n, const double* thetas, double* out) {
  if (int rc = check_plan(c, plan)) return rc;
  const FitPlan& fp = plans(c)[plan];
  unsigned nblk = ceil_div(c->npix, kFitWG);
  if (int rc = ensure_work(c, ((size_t)nblk * kNOut + (size_t)n * kNOut) * sizeof(double))) return rc;
  double* part = reinterpret_cast<double*>(c->d_work);
  double* res = part + (size_t)nblk * kNOut;
  const bool norm = normalised(fp.metric);
  for (int i = 0; i < n; ++i) {
    FitArgs a;
    if (int rc = pack_fit(c, fp, thetas + 6 * (size_t)i, &a)) return rc;
    // the next pass reuses `part` only after this finish has read it (same stream)
    if (int rc = launch_fit(c, fp, a, norm ? kModeStats : kModeValue, nullptr, part, res + (size_t)i * kNOut))
      return rc;
  }
  std::vector<double> host((size_t)n * kNOut);
  HGX_CHECK_HIP(c, hipMemcpyAsync(host.data(), res, host.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i) {
    double k[3];
    out[i] = norm ? metric_from_sums(fp.metric, &host[(size_t)i * kNOut], k) : host[(size_t)i * kNOut];
  }
  return HGX_OK;
}

// Projected L-BFGS (memory 6) over the free parameters within [lo, hi], in coordinates scaled by
// each parameter's range (so a tilt in radians and a lattice constant in nm are comparable): two-loop
// recursion on the projected gradient, backtracking (Armijo, c1 = 1e-4) along the projected path,
// each trial point one value-and-gradient pass. Stops when the projected gradient's largest scaled
// component falls below tol * |f|, or when a line search cannot decrease f.
static int run_lbfgs(Context* c, int plan, double theta[6], const int freep[6], const double lo[6],
                     const double hi[6], int iters, double tol, double* f_out, int* done) {
  const int M = 6;
  double sc[6], x[6], g[6], f;
  for (int j = 0; j < 6; ++j) {
    x[j] = std::min(std::max(theta[j], lo[j]), hi[j]);
    const double r = hi[j] - lo[j];
    sc[j] = (std::isfinite(r) && r > 0) ? r : std::max(std::fabs(x[j]), 1.0);
  }
  if (int rc = run_fit(c, plan, x, HGX_PREC_FP64, true, &f, g, nullptr)) return rc;
  std::vector<std::array<double, 6>> S, Y;  // scaled coordinates
  std::vector<double> rho;
  // projected gradient in scaled coordinates: d f / d(x_j / sc_j)
  auto proj_grad = [&](const double* xx, const double* gg, double* pg) {
    for (int j = 0; j < 6; ++j) {
      pg[j] = freep[j] ? gg[j] * sc[j] : 0.0;
      if ((xx[j] <= lo[j] && pg[j] > 0) || (xx[j] >= hi[j] && pg[j] < 0)) pg[j] = 0.0;
    }
  };
  auto clip = [&](double* d, const double* pg, double frac) {  // largest free move = frac (scaled)
    double dm = 0;
    for (int j = 0; j < 6; ++j)
      if (freep[j]) dm = std::max(dm, std::fabs(d[j]) / frac);
    if (dm > 1)
      for (int j = 0; j < 6; ++j) d[j] /= dm;
  };
  int it = 0;
  for (; it < iters; ++it) {
    double pg[6];
    proj_grad(x, g, pg);
    double gmax = 0.0;
    for (int j = 0; j < 6; ++j) gmax = std::max(gmax, std::fabs(pg[j]));
    if (gmax <= tol * std::max(std::fabs(f), 1e-300)) break;
    // two-loop recursion (scaled coordinates)
    double d[6], al[M];
    for (int j = 0; j < 6; ++j) d[j] = -pg[j];
    for (int k = (int)S.size() - 1; k >= 0; --k) {
      double sd = 0;
      for (int j = 0; j < 6; ++j) sd += S[k][j] * d[j];
      al[k] = rho[k] * sd;
      for (int j = 0; j < 6; ++j) d[j] -= al[k] * Y[k][j];
    }
    if (!S.empty()) {
      const auto& s = S.back();
      const auto& y = Y.back();
      double sy = 0, yy = 0;
      for (int j = 0; j < 6; ++j) {
        sy += s[j] * y[j];
        yy += y[j] * y[j];
      }
      for (int j = 0; j < 6; ++j) d[j] *= sy / yy;
    } else {
      double gm = 0;
      for (int j = 0; j < 6; ++j) gm = std::max(gm, std::fabs(d[j]));
      if (gm > 0)
        for (int j = 0; j < 6; ++j) d[j] *= 0.01 / gm;  // first step: 1 % of the range
    }
    for (size_t k = 0; k < S.size(); ++k) {
      double yd = 0;
      for (int j = 0; j < 6; ++j) yd += Y[k][j] * d[j];
      double be = rho[k] * yd;
      for (int j = 0; j < 6; ++j) d[j] += S[k][j] * (al[k] - be);
    }
    for (int j = 0; j < 6; ++j)
      if (!freep[j]) d[j] = 0.0;
    double gd = 0;
    for (int j = 0; j < 6; ++j) gd += pg[j] * d[j];
    if (!(gd < 0)) {  // not a descent direction: restart from the projected steepest descent
      S.clear();
      Y.clear();
      rho.clear();
      double gm = 0;
      for (int j = 0; j < 6; ++j) {
        d[j] = -pg[j];
        gm = std::max(gm, std::fabs(d[j]));
      }
      for (int j = 0; j < 6; ++j) d[j] *= gm > 0 ? 0.01 / gm : 0.0;
    }
    clip(d, pg, 0.25);  // no trial moves a parameter by more than a quarter of its range
    double step = 1.0, xn[6], gn[6], fn = f;
    bool ok = false;
    for (int ls = 0; ls < 30; ++ls) {
      double dec = 0;
      for (int j = 0; j < 6; ++j) {
        xn[j] = std::min(std::max(x[j] + step * d[j] * sc[j], lo[j]), hi[j]);
        dec += pg[j] * (xn[j] - x[j]) / sc[j];
      }
      int rc = run_fit(c, plan, xn, HGX_PREC_FP64, true, &fn, gn, nullptr);
      if (rc == HGX_OK && std::isfinite(fn) && fn <= f + 1e-4 * dec) {
        ok = true;
        break;
      }
      if (rc != HGX_OK && rc != HGX_ERR_INVALID_ARG) return rc;  // INVALID_ARG: outside the model's domain
      step *= 0.5;
    }
    if (!ok) break;
    std::array<double, 6> s, y;
    double sy = 0;
    for (int j = 0; j < 6; ++j) {
      s[j] = (xn[j] - x[j]) / sc[j];
      y[j] = freep[j] ? (gn[j] - g[j]) * sc[j] : 0.0;
      sy += s[j] * y[j];
    }
    if (sy > 1e-300) {
      if ((int)S.size() == M) {
        S.erase(S.begin());
        Y.erase(Y.begin());
        rho.erase(rho.begin());
      }
      S.push_back(s);
      Y.push_back(y);
      rho.push_back(1.0 / sy);
    }
    for (int j = 0; j < 6; ++j) {
      x[j] = xn[j];
      g[j] = gn[j];
    }
    f = fn;
  }
  for (int j = 0; j < 6; ++j) theta[j] = x[j];
  if (f_out) *f_out = f;
  if (done) *done = it;
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

extern "C" int hgx_fit_plan_metric(hgx_ctx* h, int plan, int metric) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  auto& P = plans(c);
  if (plan < 0 || plan >= (int)P.size() || !P[plan].live) return set_error(c, HGX_ERR_BAD_HANDLE, "fit: unknown plan");
  if (metric < HGX_METRIC_CHI2 || metric > HGX_METRIC_SQRT_C_L2)
    return set_error(c, HGX_ERR_INVALID_ARG, "fit_plan_metric: unknown metric %d", metric);
  P[plan].metric = metric;
  return HGX_OK;
}

extern "C" int hgx_chi2_fwd_many(hgx_ctx* h, int plan, int n, const double* thetas, int precision, double* out) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (n < 1 || !thetas || !out) return set_error(c, HGX_ERR_INVALID_ARG, "chi2_fwd_many: bad arguments");
  if (precision != HGX_PREC_FP64) return set_error(c, HGX_ERR_UNSUPPORTED, "fit: FP64 only");
  if (int rc = set_device(c)) return rc;
  return run_fit_many(c, plan, n, thetas, out);
}

extern "C" int hgx_fit_lbfgs(hgx_ctx* h, int plan, double theta[6], const int free_params[6], const double lo[6],
                             const double hi[6], int iterations, double tol, double* objective, int* iterations_done) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!theta || !free_params || !lo || !hi || iterations < 0 || !(tol >= 0))
    return set_error(c, HGX_ERR_INVALID_ARG, "fit_lbfgs: bad arguments");
  for (int j = 0; j < 6; ++j)
    if (!(lo[j] <= hi[j])) return set_error(c, HGX_ERR_INVALID_ARG, "fit_lbfgs: lo > hi for parameter %d", j);
  if (int rc = set_device(c)) return rc;
  return run_lbfgs(c, plan, theta, free_params, lo, hi, iterations, tol, objective, iterations_done);
}
