// ensemble.hip -- fused intensity of a grain ensemble whose unit cell holds any analytic elements.
//
// For every pixel p and grain g:
//     I[p] (+)= sum_g | w_g sum_b fc_b S_g(R_g q_b) sum_e dn2_e F_e(M_ge q_b) sum_l e^{i (M_ge q_b).t_l} |^2
// with M_ge = R_g R_e the element's rotation inside the grain (src/sim/hipgisaxs_main.cpp:986): the
// grain loop, unit-cell element and location sums, structure factor and DWBA sum of run_gisaxs
// (:925-1100) in one pass per pixel, without the complex [nbranch npix] form-factor and structure-
// factor arrays the per-grain calls write and re-read. FP64 throughout, with the device functions of
// the per-shape and structure-factor kernels (analytic_shapes.hpp, lattice_sf.hpp), so it matches
// hgx_ff_analytic + hgx_caxpy + hgx_sf_lattice + hgx_dwba_intensity to round-off.
//
// FP32C (real q; sphere, cylinder, box and cube cells): q rotated, lattice, location and sinc phases
// formed and reduced to turns in FP64, the form-factor node sums and Dirichlet kernels in FP32 with
// the turn-based hardware sin/cos (Taylor series at small arguments), amplitudes summed and squared
// in FP64 -- the scheme of the FP32C lattice kernel (lattice.hip).
//
// Work split: blockIdx.x tiles the pixels (256 per workgroup), blockIdx.y a contiguous range of grains;
// each (x, y) writes its partial grain sum, and a second kernel adds the y partials in a fixed order
// (deterministic). The model (nodes, matrices, locations, grain data) sits in the context workspace.
#include <algorithm>
#include <cmath>
#include <vector>

#include "analytic_shapes.hpp"
#include "analytic32.hpp"
#include "fp32_math.hpp"
#include "lattice_sf.hpp"

namespace hgx {

struct EnsElem {
  Nodes P;
  int shape;
  int loc0, nloc;  // range in the location table
  double dn2r, dn2i;
};

struct EnsGrain {
  double rot[9];     // R_g, for the structure factor
  double center[3];  // grain centre
  double w;          // grain weight
};

struct QViewE {
  const double* __restrict__ qx;
  const double* __restrict__ qy;
  const double* __restrict__ qz;
  int64_t npix;
};

__device__ inline cd ff_any(int shape, const cd mq[3], const Nodes& P) {
  switch (shape) {
    case HGX_SHAPE_SPHERE: return shape_ff<HGX_SHAPE_SPHERE>(mq, P);
    case HGX_SHAPE_CYLINDER: return shape_ff<HGX_SHAPE_CYLINDER>(mq, P);
    case HGX_SHAPE_BOX: return shape_ff<HGX_SHAPE_BOX>(mq, P);
    case HGX_SHAPE_CUBE: return shape_ff<HGX_SHAPE_CUBE>(mq, P);
    case HGX_SHAPE_PRISM3: return shape_ff<HGX_SHAPE_PRISM3>(mq, P);
    case HGX_SHAPE_PRISM6: return shape_ff<HGX_SHAPE_PRISM6>(mq, P);
    default: return shape_ff<HGX_SHAPE_PYRAMID>(mq, P);
  }
}

__device__ inline void rotate_m(const double* __restrict__ m, double qx, double qy, cd qz, cd mq[3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    mq[j].r = m[3 * j] * qx + m[3 * j + 1] * qy + m[3 * j + 2] * qz.r;
    mq[j].i = m[3 * j + 2] * qz.i;
  }
}

// The unit cell's form-factor sum for grain o at one q-point:
//     sum_e dn2_e F_e(M_oe q) sum_l exp(i (M_oe q).t_l).
template <int SHAPE>
__device__ inline cd cell64(int o, double qx, double qy, cd qz, const EnsElem* __restrict__ E, int nelem,
                            const double* __restrict__ loc, const double* __restrict__ M) {
  cd ff = cd{0.0, 0.0};
  for (int e = 0; e < nelem; ++e) {
    cd mq[3];
    rotate_m(M + 9 * ((int64_t)o * nelem + e), qx, qy, qz, mq);
    const cd F = SHAPE >= 0 ? shape_ff<(SHAPE >= 0 ? SHAPE : 0)>(mq, E[e].P) : ff_any(E[e].shape, mq, E[e].P);
    const cd dn2 = cd{E[e].dn2r, E[e].dn2i};
    for (int l = E[e].loc0; l < E[e].loc0 + E[e].nloc; ++l) {
      const double* t = loc + 3 * l;
      ff = ff + dn2 * (F * cexpi_d(mq[0] * t[0] + mq[1] * t[1] + mq[2] * t[2]));
    }
  }
  return ff;
}

// SHAPE >= 0: every element has that shape (one form-factor function inlined, fewer registers);
// SHAPE = -1: mixed cells, the shape chosen per element. HOIST: the cell's form-factor sum is the same
// for every grain (the host checked it: spheres and cylinders whose F depends on q only through q.q
// and the third row of R_g R_e, which every grain shares, with locations that every grain sees
// alike), so it is evaluated once per branch and the grain loop runs the structure factor alone.
template <int SHAPE, bool HOIST>
__global__ __launch_bounds__(256) void k_ensemble64(QViewE g, int nb, const double* __restrict__ fc,
                                                    const EnsElem* __restrict__ E, int nelem,
                                                    const double* __restrict__ loc, const double* __restrict__ M,
                                                    const EnsGrain* __restrict__ G, int ngrains, int per_y,
                                                    SFArgs sfb, double* __restrict__ part) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const int g0 = blockIdx.y * per_y, g1 = min(ngrains, g0 + per_y);
  const double qx = g.qx[p], qy = g.qy[p];
  cd qzb[4], fcb[4], ffh[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if (b >= nb) break;
    const int64_t q = p + b * g.npix;
    qzb[b] = cd{g.qz[2 * q], g.qz[2 * q + 1]};
    fcb[b] = fc ? cd{fc[2 * q], fc[2 * q + 1]} : cd{1.0, 0.0};
    if (HOIST) ffh[b] = fcb[b] * cell64<SHAPE>(g0, qx, qy, qzb[b], E, nelem, loc, M);
  }
  double acc = 0.0;
  for (int o = g0; o < g1; ++o) {
    SFArgs sf = sfb;
    sf.center[0] = G[o].center[0];
    sf.center[1] = G[o].center[1];
    sf.center[2] = G[o].center[2];
    cd amp = cd{0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (b >= nb) break;
      cd mg[3];
      rotate_m(G[o].rot, qx, qy, qzb[b], mg);
      const cd S = sf_value(mg, sf);
      if (HOIST) {
        amp = amp + ffh[b] * S;
      } else {
        cd ff = cd{0.0, 0.0};
        for (int e = 0; e < nelem; ++e) {
          cd mq[3];
          rotate_m(M + 9 * ((int64_t)o * nelem + e), qx, qy, qzb[b], mq);
          const cd F = SHAPE >= 0 ? shape_ff<(SHAPE >= 0 ? SHAPE : 0)>(mq, E[e].P) : ff_any(E[e].shape, mq, E[e].P);
          const cd dn2 = cd{E[e].dn2r, E[e].dn2i};
          for (int l = E[e].loc0; l < E[e].loc0 + E[e].nloc; ++l) {
            const double* t = loc + 3 * l;
            ff = ff + dn2 * (F * cexpi_d(mq[0] * t[0] + mq[1] * t[1] + mq[2] * t[2]));
          }
        }
        amp = amp + fcb[b] * S * ff;
      }
    }
    amp = G[o].w * amp;
    acc += amp.r * amp.r + amp.i * amp.i;
  }
  part[(int64_t)blockIdx.y * g.npix + p] = acc;
}

// ----- FP32C -------------------------------------------------------------------------------------

// FP32C form-factor sum of the unit cell for grain o (FP64 result of FP32 terms).
template <int SHAPE>
__device__ inline cd cell32(int o, double qx, double qy, double qz, const EnsElem* __restrict__ E, int nelem,
                            const double* __restrict__ loc, const double* __restrict__ M) {
  cd ff = cd{0.0, 0.0};
  for (int e = 0; e < nelem; ++e) {
    const double* Me = M + 9 * ((int64_t)o * nelem + e);
    double mq[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) mq[j] = Me[3 * j] * qx + Me[3 * j + 1] * qy + Me[3 * j + 2] * qz;
    const cd F = SHAPE >= 0 ? ff_fp32c<(SHAPE >= 0 ? SHAPE : 0)>(mq, E[e].P) : ff_fp32c_any(E[e].shape, mq, E[e].P);
    cf ph = cf{0.0f, 0.0f};
    for (int l = E[e].loc0; l < E[e].loc0 + E[e].nloc; ++l) {
      const double* t = loc + 3 * l;
      ph = ph + expi_turn(turns_of(mq[0] * t[0] + mq[1] * t[1] + mq[2] * t[2]));
    }
    ff = ff + cd{E[e].dn2r, E[e].dn2i} * (F * cd{(double)ph.r, (double)ph.i});
  }
  return ff;
}

template <int SHAPE, bool HOIST>
__global__ __launch_bounds__(256) void k_ensemble32(QViewE g, int nb, const double* __restrict__ fc,
                                                    const EnsElem* __restrict__ E, int nelem,
                                                    const double* __restrict__ loc, const double* __restrict__ M,
                                                    const EnsGrain* __restrict__ G, int ngrains, int per_y,
                                                    SFArgs sfb, double* __restrict__ part) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const int g0 = blockIdx.y * per_y, g1 = min(ngrains, g0 + per_y);
  const double qx = g.qx[p], qy = g.qy[p];
  double qzb[4];
  cd fcb[4], ffh[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if (b >= nb) break;
    const int64_t q = p + b * g.npix;
    qzb[b] = g.qz[2 * q];
    fcb[b] = fc ? cd{fc[2 * q], fc[2 * q + 1]} : cd{1.0, 0.0};
    if (HOIST) ffh[b] = fcb[b] * cell32<SHAPE>(g0, qx, qy, qzb[b], E, nelem, loc, M);
  }
  double acc = 0.0;
  for (int o = g0; o < g1; ++o) {
    SFArgs sf = sfb;
    sf.center[0] = G[o].center[0];
    sf.center[1] = G[o].center[1];
    sf.center[2] = G[o].center[2];
    cd amp = cd{0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (b >= nb) break;
      double mg[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) mg[j] = G[o].rot[3 * j] * qx + G[o].rot[3 * j + 1] * qy + G[o].rot[3 * j + 2] * qzb[b];
      const cf S = sf32e(mg, sf);
      const cd Sd = cd{(double)S.r, (double)S.i};
      if (HOIST) {
        amp = amp + ffh[b] * Sd;
      } else {
        amp = amp + fcb[b] * (Sd * cell32<SHAPE>(o, qx, qy, qzb[b], E, nelem, loc, M));
      }
    }
    amp = G[o].w * amp;
    acc += amp.r * amp.r + amp.i * amp.i;
  }
  part[(int64_t)blockIdx.y * g.npix + p] = acc;
}

__global__ __launch_bounds__(256) void k_ensemble_sum(int64_t npix, int ny, const double* __restrict__ part,
                                                      int accumulate, double* __restrict__ I) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= npix) return;
  double s = accumulate ? I[p] : 0.0;
  for (int y = 0; y < ny; ++y) s += part[(int64_t)y * npix + p];
  I[p] = s;
}

}  // namespace hgx

using namespace hgx;

static const int kEnsDims[7] = {1, 2, 3, 1, 2, 2, 4};

namespace hgx {

// Whether the cell's form-factor sum is the same for every grain: every element a sphere or a
// cylinder (F depends on q only through q.q and (M q)_z, so on the third row of M = R_g R_e), that row
// equal for every grain, and every location seen alike (M_ge^T t_l equal for every grain, so the
// phases (M q).t_l are). Rotations about z of cells at the origin, the reference's orientation
// distributions of spheres and cylinders, meet it.
static bool cell_is_grain_invariant(const hgx_ensemble_model* m, const std::vector<double>& Mm) {
  const int ne = m->nelem, ng = m->ngrains;
  auto close = [](double a, double b, double scale) { return std::fabs(a - b) <= 1e-14 * scale; };
  for (int e = 0; e < ne; ++e) {
    const int sh = m->elem[e].shape;
    if (sh != HGX_SHAPE_SPHERE && sh != HGX_SHAPE_CYLINDER) return false;
    const double* M0 = &Mm[9 * (size_t)e];
    for (int o = 1; o < ng; ++o) {
      const double* Mo = &Mm[9 * ((size_t)o * ne + e)];
      for (int k = 6; k < 9; ++k)
        if (!close(Mo[k], M0[k], 1.0)) return false;
      for (int l = 0; l < m->elem[e].nloc; ++l) {
        const double* t = m->elem[e].loc + 3 * l;
        const double tn = std::fabs(t[0]) + std::fabs(t[1]) + std::fabs(t[2]);
        for (int j = 0; j < 3; ++j) {
          const double u0 = M0[j] * t[0] + M0[3 + j] * t[1] + M0[6 + j] * t[2];
          const double uo = Mo[j] * t[0] + Mo[3 + j] * t[1] + Mo[6 + j] * t[2];
          if (!close(uo, u0, tn)) return false;
        }
      }
    }
  }
  return true;
}

static int ensemble_run(Context* c, const hgx_ensemble_model* m, int precision, const double* fc, int accumulate,
                        double* I) {
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "ensemble_intensity: no q-grid");
  if (!m || !I || m->nelem < 1 || !m->elem || m->ngrains < 1 || !m->rot)
    return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_intensity: need >= 1 element, >= 1 grain and an output");
  if (m->nelem > 64) return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_intensity: at most 64 elements");
  if (precision != HGX_PREC_FP64 && precision != HGX_PREC_FP32C)
    return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_intensity: bad precision");
  const int ne = m->nelem, ng = m->ngrains;
  if ((int64_t)ng * ne > ((int64_t)1 << 24))
    return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_intensity: grains x elements above 2^24");
  std::vector<EnsElem> E(ne);
  std::vector<double> loc;
  for (int e = 0; e < ne; ++e) {
    const hgx_ensemble_element& el = m->elem[e];
    if (el.shape < 0 || el.shape > HGX_SHAPE_PYRAMID)
      return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_intensity: element %d has bad shape %d", e, el.shape);
    if (el.nloc < 1 || !el.loc) return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_intensity: element %d needs locations", e);
    Nodes& P = E[e].P;
    for (int d = 0; d < 4; ++d) {
      P.n[d] = 0;
      if (d >= kEnsDims[el.shape]) continue;
      int n = el.params.n[d];
      if (n < 1 || n > HGX_MAX_NODES || !el.params.val[d] || !el.params.wt[d])
        return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_intensity: element %d dimension %d needs 1..%d nodes", e,
                         d, HGX_MAX_NODES);
      P.n[d] = n;
      for (int i = 0; i < n; ++i) {
        P.val[d][i] = el.params.val[d][i];
        P.wt[d][i] = el.params.wt[d][i];
      }
    }
    P.tan_tau = std::tan(el.params.tau);
    P.sin_eta = std::sin(el.params.eta);
    P.cos_eta = std::cos(el.params.eta);
    P.flags = el.params.flags;
    E[e].shape = el.shape;
    E[e].loc0 = (int)(loc.size() / 3);
    E[e].nloc = el.nloc;
    E[e].dn2r = el.dn2_re;
    E[e].dn2i = el.dn2_im;
    loc.insert(loc.end(), el.loc, el.loc + 3 * el.nloc);
  }
  int common = E[0].shape;
  for (int e = 1; e < ne; ++e)
    if (E[e].shape != common) common = -1;
  if (precision == HGX_PREC_FP32C && !c->qz_real)
    return set_error(c, HGX_ERR_UNSUPPORTED, "ensemble_intensity: FP32C needs real q");
  // M[g][e] = R_g R_e with the summation order of the reference's RotMatrix_t product
  std::vector<double> Mm(9 * (size_t)ng * ne);
  std::vector<EnsGrain> G(ng);
  for (int o = 0; o < ng; ++o) {
    const double* Rg = m->rot + 9 * (size_t)o;
    for (int k = 0; k < 9; ++k) G[o].rot[k] = Rg[k];
    for (int k = 0; k < 3; ++k) G[o].center[k] = m->center ? m->center[3 * (size_t)o + k] : 0.0;
    G[o].w = m->weight ? m->weight[o] : 1.0;
    for (int e = 0; e < ne; ++e) {
      const double* Re = m->elem[e].rot;
      double* out = &Mm[9 * ((size_t)o * ne + e)];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double s = 0;
          for (int k = 0; k < 3; ++k) s += Rg[3 * i + k] * Re[3 * k + j];
          out[3 * i + j] = s;
        }
    }
  }
  const bool hoist = cell_is_grain_invariant(m, Mm);
  SFArgs sfb;
  for (int k = 0; k < 3; ++k) {
    sfb.la[k] = m->la[k];
    sfb.lb[k] = m->lb[k];
    sfb.lc[k] = m->lc[k];
    sfb.N[k] = m->repet[k] < 1 ? 1 : m->repet[k];  // src/sf/sf.cpp:165-167
    sfb.center[k] = 0.0;
  }
  sfb.dw = 0.5 * m->dw_sigma * m->dw_sigma;
  sfb.compat_lb2 = (m->sf_flags & HGX_SF_COMPAT_LB2) != 0;
  if (int rc = set_device(c)) return rc;
  // grains split over blockIdx.y so that a launch has at least ~2048 workgroups
  const unsigned nbx = ceil_div(c->npix, 256);
  int ny = std::max(1, std::min(ng, (int)((2048 + nbx - 1) / nbx)));
  const int per_y = (ng + ny - 1) / ny;
  ny = (ng + per_y - 1) / per_y;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t b_part = al((size_t)ny * c->npix * sizeof(double)), b_E = al(ne * sizeof(EnsElem)),
               b_loc = al(loc.size() * sizeof(double)), b_M = al(Mm.size() * sizeof(double)),
               b_G = al(ng * sizeof(EnsGrain));
  if (int rc = ensure_work(c, b_part + b_E + b_loc + b_M + b_G)) return rc;
  char* w = static_cast<char*>(c->d_work);
  double* part = reinterpret_cast<double*>(w);
  EnsElem* dE = reinterpret_cast<EnsElem*>(w + b_part);
  double* dloc = reinterpret_cast<double*>(w + b_part + b_E);
  double* dM = reinterpret_cast<double*>(w + b_part + b_E + b_loc);
  EnsGrain* dG = reinterpret_cast<EnsGrain*>(w + b_part + b_E + b_loc + b_M);
  {
    const void* src[4] = {E.data(), loc.data(), Mm.data(), G.data()};
    const size_t nb[4] = {ne * sizeof(EnsElem), loc.size() * sizeof(double), Mm.size() * sizeof(double),
                          ng * sizeof(EnsGrain)};
    if (int rc = stage_to_device(c, 4, src, nb, reinterpret_cast<char*>(dE))) return rc;
  }
  QViewE g{c->d_qx, c->d_qy, c->d_qz, c->npix};
  if (precision == HGX_PREC_FP32C) {
    LaunchTimer lt(c, true);
    const dim3 grid(nbx, ny);
#define HGX_ENS32_LAUNCH(S)                                                                                    \
  if (hoist)                                                                                                   \
    k_ensemble32<S, true><<<grid, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, ne, dloc, dM, dG, ng, per_y, sfb, \
                                                       part);                                                  \
  else                                                                                                         \
    k_ensemble32<S, false><<<grid, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, ne, dloc, dM, dG, ng, per_y, sfb, part)
    switch (common) {
      case HGX_SHAPE_SPHERE: HGX_ENS32_LAUNCH(HGX_SHAPE_SPHERE); break;
      case HGX_SHAPE_CYLINDER: HGX_ENS32_LAUNCH(HGX_SHAPE_CYLINDER); break;
      case HGX_SHAPE_BOX: HGX_ENS32_LAUNCH(HGX_SHAPE_BOX); break;
      case HGX_SHAPE_CUBE: HGX_ENS32_LAUNCH(HGX_SHAPE_CUBE); break;
      case HGX_SHAPE_PRISM3: HGX_ENS32_LAUNCH(HGX_SHAPE_PRISM3); break;
      case HGX_SHAPE_PRISM6: HGX_ENS32_LAUNCH(HGX_SHAPE_PRISM6); break;
      case HGX_SHAPE_PYRAMID: HGX_ENS32_LAUNCH(HGX_SHAPE_PYRAMID); break;
      default: HGX_ENS32_LAUNCH(-1); break;
    }
#undef HGX_ENS32_LAUNCH
    HGX_CHECK_LAUNCH(c);
  } else {
    LaunchTimer lt(c, true);
    const dim3 grid(nbx, ny);
#define HGX_ENS_LAUNCH(S)                                                                                      \
  if (hoist)                                                                                                   \
    k_ensemble64<S, true><<<grid, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, ne, dloc, dM, dG, ng, per_y, sfb, \
                                                       part);                                                  \
  else                                                                                                         \
    k_ensemble64<S, false><<<grid, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, ne, dloc, dM, dG, ng, per_y, sfb, part)
    switch (common) {
      case HGX_SHAPE_SPHERE: HGX_ENS_LAUNCH(HGX_SHAPE_SPHERE); break;
      case HGX_SHAPE_CYLINDER: HGX_ENS_LAUNCH(HGX_SHAPE_CYLINDER); break;
      case HGX_SHAPE_BOX: HGX_ENS_LAUNCH(HGX_SHAPE_BOX); break;
      case HGX_SHAPE_CUBE: HGX_ENS_LAUNCH(HGX_SHAPE_CUBE); break;
      case HGX_SHAPE_PRISM3: HGX_ENS_LAUNCH(HGX_SHAPE_PRISM3); break;
      case HGX_SHAPE_PRISM6: HGX_ENS_LAUNCH(HGX_SHAPE_PRISM6); break;
      case HGX_SHAPE_PYRAMID: HGX_ENS_LAUNCH(HGX_SHAPE_PYRAMID); break;
      default: HGX_ENS_LAUNCH(-1); break;
    }
#undef HGX_ENS_LAUNCH
    HGX_CHECK_LAUNCH(c);
  }
  {
    LaunchTimer lt(c, false);
    k_ensemble_sum<<<nbx, 256, 0, c->stream>>>(c->npix, ny, part, accumulate, I);
    HGX_CHECK_LAUNCH(c);
  }
  return HGX_OK;
}

}  // namespace hgx

extern "C" int hgx_ensemble_intensity(hgx_ctx* h, const hgx_ensemble_model* m, int precision, const double* fc,
                                      int accumulate, double* I) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  return ensemble_run(c, m, precision, fc, accumulate, I);
}

// A lattice of one particle shape at the cell origin: the ensemble path with one element (identity
// rotation, one location at the origin) and every grain at the model's centre.
extern "C" int hgx_lattice_intensity(hgx_ctx* h, const hgx_lattice_model* m, int precision, const double* fc,
                                     int accumulate, double* I) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!m || !I || !m->rot || m->ngrains < 1)
    return set_error(c, HGX_ERR_INVALID_ARG, "lattice_intensity: need a model with >= 1 grain and an output");
  if (m->shape != HGX_SHAPE_SPHERE && m->shape != HGX_SHAPE_CYLINDER)
    return set_error(c, HGX_ERR_UNSUPPORTED, "lattice: shape %d not supported (sphere, cylinder)", m->shape);
  if (m->ngrains > (1 << 24)) return set_error(c, HGX_ERR_INVALID_ARG, "lattice_intensity: too many grains");
  static const double origin[3] = {0.0, 0.0, 0.0};
  hgx_ensemble_element el{};
  el.shape = m->shape;
  el.params = m->params;
  for (int k = 0; k < 9; ++k) el.rot[k] = (k % 4 == 0) ? 1.0 : 0.0;
  el.nloc = 1;
  el.loc = origin;
  el.dn2_re = m->dn2_re;
  el.dn2_im = m->dn2_im;
  std::vector<double> centers(3 * (size_t)m->ngrains);
  for (int o = 0; o < m->ngrains; ++o)
    for (int k = 0; k < 3; ++k) centers[3 * (size_t)o + k] = m->center[k];
  hgx_ensemble_model em{};
  em.nelem = 1;
  em.elem = &el;
  for (int k = 0; k < 3; ++k) {
    em.la[k] = m->la[k];
    em.lb[k] = m->lb[k];
    em.lc[k] = m->lc[k];
    em.repet[k] = m->repet[k];
  }
  em.dw_sigma = m->dw_sigma;
  em.sf_flags = m->sf_flags;
  em.ngrains = m->ngrains;
  em.rot = m->rot;
  em.center = centers.data();
  em.weight = m->weight;
  return ensemble_run(c, &em, precision, fc, accumulate, I);
}
