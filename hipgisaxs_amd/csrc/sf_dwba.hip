// sf_dwba.hip -- lattice structure factor, Parratt/Fresnel coefficients and the DWBA combination.
#include <cmath>

#include "analytic32.hpp"
#include "lattice_sf.hpp"

namespace hgx {

constexpr int kMaxLayers = 32;

__global__ __launch_bounds__(256) void k_sf_lattice(const double* __restrict__ qx, const double* __restrict__ qy,
                                                    const double* __restrict__ qz, int64_t npix, int64_t nq, Rot R,
                                                    SFArgs a, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nq) return;
  int64_t j = i % npix;
  cd mq[3];
  rotate_q(R, qx[j], qy[j], cd{qz[2 * i], qz[2 * i + 1]}, mq);
  cd s = sf_value(mq, a);
  out[2 * i] = s.r;
  out[2 * i + 1] = s.i;
}

// FP32C: lattice and centre phases formed and reduced to turns in FP64, Dirichlet kernels in FP32
// (Taylor sines next to the Bragg peaks), real q only.
__global__ __launch_bounds__(256) void k_sf_lattice32(const double* __restrict__ qx, const double* __restrict__ qy,
                                                      const double* __restrict__ qz, int64_t npix, int64_t nq, Rot R,
                                                      SFArgs a, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nq) return;
  int64_t j = i % npix;
  const double x = qx[j], y = qy[j], z = qz[2 * i];
  double m[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) m[k] = R.m[3 * k] * x + R.m[3 * k + 1] * y + R.m[3 * k + 2] * z;
  const cf s = sf32e(m, a);
  out[2 * i] = (double)s.r;
  out[2 * i + 1] = (double)s.i;
}

// ----- Parratt recursion (src/model/multilayer.cpp:65-108) --------------------------------------

struct Stack {
  int NL;  // including the vacuum
  int order;
  double n2r[kMaxLayers + 1], n2i[kMaxLayers + 1];
  double z[kMaxLayers + 1];
};

__device__ void parratt(const Stack& S, double alpha, double k0, cd* Tout, cd* Rout) {
  double sina = sin(alpha);
  cd T_next = cd{1.0, 0.0}, R_next = cd{0.0, 0.0};
  cd kz_next = -k0 * csqrt_d(cd{sina * sina - S.n2r[S.NL - 1], -S.n2i[S.NL - 1]});
  cd T_ord = S.order == S.NL - 1 ? T_next : cd{0, 0}, R_ord = S.order == S.NL - 1 ? R_next : cd{0, 0};
  cd T0 = T_next;
  for (int i = S.NL - 2; i > -1; --i) {
    cd kz = -k0 * csqrt_d(cd{sina * sina - S.n2r[i], -S.n2i[i]});
    cd inv2k = cdiv(cd{1.0, 0.0}, 2.0 * kz);
    cd pij = (kz + kz_next) * inv2k;
    cd mij = (kz - kz_next) * inv2k;
    double z = S.z[i];
    cd exp_p = cexpi_d(-(z * (kz_next + kz)));
    cd exp_m = cexpi_d(-(z * (kz_next - kz)));
    cd a00 = pij * exp_m, a01 = mij * conj(exp_p);
    cd a10 = mij * exp_p, a11 = pij * conj(exp_m);
    cd T = a00 * T_next + a01 * R_next;
    cd Rr = a10 * T_next + a11 * R_next;
    T_next = T;
    R_next = Rr;
    kz_next = kz;
    if (i == S.order) {
      T_ord = T;
      R_ord = Rr;
    }
    T0 = T;
  }
  *Tout = cdiv(T_ord, T0);
  *Rout = cdiv(R_ord, T0);
}

// rows [0, nrow): Tf, Rf of alpha_f (zero for alpha_f <= 0); row nrow: Ti, Ri of alpha_i.
__global__ void k_fresnel_rows(Stack S, int nrow, const double* __restrict__ alpha_f, double alpha_i, double k0,
                               double* __restrict__ tr) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > nrow) return;
  cd T = cd{0.0, 0.0}, Rr = cd{0.0, 0.0};
  double a = r == nrow ? alpha_i : alpha_f[r];
  if (r == nrow || a > 0) parratt(S, a, k0, &T, &Rr);
  tr[4 * r] = T.r;
  tr[4 * r + 1] = T.i;
  tr[4 * r + 2] = Rr.r;
  tr[4 * r + 3] = Rr.i;
}

// fc = [Ti conj Tf, Ti conj Rf, Ri conj Tf, Ri conj Rf] (multilayer.cpp:146-153)
__global__ __launch_bounds__(256) void k_fresnel_fill(int64_t npix, int ncol, int nrow, const double* __restrict__ tr,
                                                      double* __restrict__ fc) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= npix) return;
  int64_t j = i / ncol;
  cd Tf = cd{tr[4 * j], tr[4 * j + 1]}, Rf = cd{tr[4 * j + 2], tr[4 * j + 3]};
  cd Ti = cd{tr[4 * nrow], tr[4 * nrow + 1]}, Ri = cd{tr[4 * nrow + 2], tr[4 * nrow + 3]};
  cd v[4] = {Ti * conj(Tf), Ti * conj(Rf), Ri * conj(Tf), Ri * conj(Rf)};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    fc[2 * (i + k * npix)] = v[k].r;
    fc[2 * (i + k * npix) + 1] = v[k].i;
  }
}

// ----- DWBA combination (src/sim/hipgisaxs_main.cpp:1084-1105) ----------------------------------

__global__ __launch_bounds__(256) void k_dwba(int64_t npix, int nbranch, const double* __restrict__ fc,
                                              const double* __restrict__ sf, const double* __restrict__ ff,
                                              double weight, int accumulate, double* __restrict__ I) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= npix) return;
  cd t = cd{0.0, 0.0};
  for (int k = 0; k < nbranch; ++k) {
    int64_t idx = i + k * npix;
    cd f = fc ? cd{fc[2 * idx], fc[2 * idx + 1]} : cd{1.0, 0.0};
    cd s = sf ? cd{sf[2 * idx], sf[2 * idx + 1]} : cd{1.0, 0.0};
    t = t + f * s * cd{ff[2 * idx], ff[2 * idx + 1]};
  }
  t = weight * t;
  double v = t.r * t.r + t.i * t.i;
  I[i] = accumulate ? I[i] + v : v;
}

}  // namespace hgx

using namespace hgx;

extern "C" int hgx_sf_lattice(hgx_ctx* h, const double la[3], const double lb[3], const double lc[3],
                              const double repet[3], const double center[3], double dw_sigma, int flags,
                              const double rot[9], double* sf) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "sf_lattice: no q-grid");
  if (!la || !lb || !lc || !repet || !sf) return set_error(c, HGX_ERR_INVALID_ARG, "sf_lattice: null argument");
  SFArgs a;
  for (int k = 0; k < 3; ++k) {
    a.la[k] = la[k];
    a.lb[k] = lb[k];
    a.lc[k] = lc[k];
    a.N[k] = repet[k] < 1 ? 1 : repet[k];  // src/sf/sf.cpp:165-167
    a.center[k] = center ? center[k] : 0.0;
  }
  a.dw = 0.5 * dw_sigma * dw_sigma;
  a.compat_lb2 = (flags & HGX_SF_COMPAT_LB2) != 0;
  if (flags & ~(HGX_SF_COMPAT_LB2 | HGX_SF_FP32C)) return set_error(c, HGX_ERR_INVALID_ARG, "sf_lattice: bad flags");
  if ((flags & HGX_SF_FP32C) && !c->qz_real)
    return set_error(c, HGX_ERR_UNSUPPORTED, "sf_lattice: FP32C needs real q (a particle in a non-absorbing layer)");
  if (int rc = set_device(c)) return rc;
  LaunchTimer lt(c, true);
  if (flags & HGX_SF_FP32C)
    k_sf_lattice32<<<ceil_div(c->nq, 256), 256, 0, c->stream>>>(c->d_qx, c->d_qy, c->d_qz, c->npix, c->nq,
                                                                make_rot(rot), a, sf);
  else
    k_sf_lattice<<<ceil_div(c->nq, 256), 256, 0, c->stream>>>(c->d_qx, c->d_qy, c->d_qz, c->npix, c->nq,
                                                              make_rot(rot), a, sf);
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

extern "C" int hgx_fresnel(hgx_ctx* h, int nl, const double* delta, const double* beta, const double* thickness,
                           int order, double k0, double alpha_i, double* fc) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "fresnel: no q-grid");
  if (!c->have_alpha) return set_error(c, HGX_ERR_INVALID_ARG, "fresnel: q-grid has no alpha_f rows");
  if (nl < 1 || nl > kMaxLayers || !delta || !beta || !thickness || !fc)
    return set_error(c, HGX_ERR_INVALID_ARG, "fresnel: need 1..%d layers", kMaxLayers);
  if (order < 0 || order > nl - 1)  // src/model/multilayer.cpp:111-116 rejects the substrate
    return set_error(c, HGX_ERR_INVALID_ARG, "fresnel: particle layer order %d outside [0, %d]", order, nl - 1);
  Stack S;
  S.NL = nl + 1;
  S.order = order;
  S.n2r[0] = S.n2i[0] = 0.0;
  S.z[0] = 0.0;
  double z = 0.0;
  for (int i = 1; i <= nl; ++i) {
    S.n2r[i] = 2 * delta[i - 1];  // RefractiveIndex::one_minus_n2 (include/model/common.hpp:36)
    S.n2i[i] = 2 * beta[i - 1];
    z -= thickness[i - 1];
    S.z[i] = z;
  }
  if (int rc = set_device(c)) return rc;
  if (int rc = ensure_work(c, (size_t)(c->nrow + 1) * 4 * sizeof(double))) return rc;
  double* tr = reinterpret_cast<double*>(c->d_work);
  k_fresnel_rows<<<ceil_div(c->nrow + 1, 64), 64, 0, c->stream>>>(S, c->nrow, c->d_alpha, alpha_i, k0, tr);
  HGX_CHECK_LAUNCH(c);
  k_fresnel_fill<<<ceil_div(c->npix, 256), 256, 0, c->stream>>>(c->npix, c->ncol, c->nrow, tr, fc);
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

extern "C" int hgx_dwba_intensity(hgx_ctx* h, const double* fc, const double* sf, const double* ff, double weight,
                                  int accumulate, double* I) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "dwba_intensity: no q-grid");
  if (!ff || !I) return set_error(c, HGX_ERR_INVALID_ARG, "dwba_intensity: null argument");
  if (int rc = set_device(c)) return rc;
  LaunchTimer lt(c, false);
  k_dwba<<<ceil_div(c->npix, 256), 256, 0, c->stream>>>(c->npix, c->nbranch, fc, sf, ff, weight, accumulate, I);
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}
