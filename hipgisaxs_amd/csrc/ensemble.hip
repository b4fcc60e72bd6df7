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
#include <cstring>
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

// One grain's structure-factor phases in the q frame. The lattice phase of direction k at q is
// (R_g q).l_k = q.(R_g^T l_k), so the rotated lattice vectors (and R_g^T c_g for the grain-centre
// phase) replace the rotation of every q-point; a rotated vector with no z component gives the same
// phase on every DWBA branch, which differ only in q_z, so its factor is taken once per grain.
struct EnsGrain {
  double L[4][3];  // R_g^T l_a, R_g^T l_b, R_g^T l_c, R_g^T c_g, times `scale` (1, or 1/2pi for turns)
  double off[4];   // constant phases (the reference's lb[2]*lb[2] in place of mq[2]*lb[2], :212), scaled
  double rot[9];   // R_g, for the Debye-Waller factor when some R_g is not orthogonal
  double w;        // grain weight
  int flat;        // bit k: L[k][2] == 0 (the factor is the same on every branch)
};

// The structure-factor constants shared by every grain: Dirichlet orders, Debye-Waller sigma^2/2 and
// whether |R_g q| = |q| for every grain (the factor then depends on the branch only).
struct EnsSF {
  double N[3];
  double dw;
  int dw_invariant;
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

// Per-branch structure factor of grain gr at q = (qx, qy, qz_b), FP64, from the grain's table:
// the branch-independent factors (flat directions) are passed in as Sflat.
__device__ inline cd sf_branch64(const EnsGrain& gr, const double u[4], cd Sflat, cd qz, const EnsSF& sf) {
  cd S = Sflat;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (!(gr.flat & (1 << k))) S = S * dirichlet(cd{u[k] + qz.r * gr.L[k][2], qz.i * gr.L[k][2]}, sf.N[k]);
  if (!(gr.flat & 8)) S = S * cexpi_d(cd{u[3] + qz.r * gr.L[3][2], qz.i * gr.L[3][2]});
  return S;
}

// |R_g q|^2 (the Debye-Waller argument) for a grain whose rotation is not orthogonal
__device__ inline double rot_norm2(const double* __restrict__ R, double qx, double qy, cd qz) {
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double re = R[3 * j] * qx + R[3 * j + 1] * qy + R[3 * j + 2] * qz.r, im = R[3 * j + 2] * qz.i;
    s += re * re + im * im;
  }
  return s;
}

// D_N of a real phase (the real-argument case of dirichlet() in lattice_sf.hpp: the same wrap into
// (-pi, pi] and the same limit N where the denominator vanishes), and e^{iu} by one sincos
// D_N(x) = sin(N x / 2) / sin(x / 2) for real x. For an integer N up to 48 (a lattice's repetition
// count, the same on every lane) it is U_{N-1}(cos(x / 2)), the Chebyshev polynomial of the second
// kind, by its three-term recurrence U_{k+1} = 2 c U_k - U_{k-1}: one cosine and N - 2 multiply-adds
// in place of two sines and a division, with the peak value N exact and no 0/0; larger N keep the
// sines (the recurrence would cost more).
__device__ inline double dirichlet_real(double x, double N) {
  const double w = x - 2.0 * kPiD * rint(x * (0.5 / kPiD));
  const double h = 0.5 * w;
  if (N <= 48.0 && N == rint(N)) {
    const int n = (int)N;
    if (n <= 1) return 1.0;
    const double c2 = 2.0 * cos(h);
    double a = 1.0, b = c2;  // U_0, U_1
    for (int k = 2; k < n; ++k) {
      const double t = c2 * b - a;
      a = b;
      b = t;
    }
    return b;
  }
  const double den = sin(h);
  if (fabs(den) < 1e-300) return N;
  return sin(N * h) / den;
}

__device__ inline cd expi_real(double u) {
  double sn, cs;
  sincos(u, &sn, &cs);
  return cd{cs, sn};
}

// The branch factors every grain shares: the Fresnel coefficient and, when every R_g is orthogonal,
// exp(-sigma^2/2 |R q|^2) = exp(-sigma^2/2 |q|^2).
__device__ inline cd branch_factor64(const double* __restrict__ fc, int64_t q, double qx, double qy, cd qz,
                                     const EnsSF& sf) {
  cd f = fc ? cd{fc[2 * q], fc[2 * q + 1]} : cd{1.0, 0.0};
  if (sf.dw != 0.0 && sf.dw_invariant) f = exp(-sf.dw * (qx * qx + qy * qy + qz.r * qz.r + qz.i * qz.i)) * f;
  return f;
}

// Fused kernel, the cell's form-factor sum per grain. SHAPE >= 0: every element has that shape (one
// form-factor function inlined, fewer registers); SHAPE = -1: mixed cells, the shape chosen per element.
template <int SHAPE>
__global__ __launch_bounds__(256) void k_ensemble64(QViewE g, int nb, const double* __restrict__ fc,
                                                    const EnsElem* __restrict__ E, int nelem,
                                                    const double* __restrict__ loc, const double* __restrict__ M,
                                                    const EnsGrain* __restrict__ G, int ngrains, int per_y,
                                                    EnsSF sf, double* __restrict__ part) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const int g0 = blockIdx.y * per_y, g1 = min(ngrains, g0 + per_y);
  const double qx = g.qx[p], qy = g.qy[p];
  cd qzb[4], fcb[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if (b >= nb) break;
    const int64_t q = p + b * g.npix;
    qzb[b] = cd{g.qz[2 * q], g.qz[2 * q + 1]};
    fcb[b] = branch_factor64(fc, q, qx, qy, qzb[b], sf);
  }
  double acc = 0.0;
  for (int o = g0; o < g1; ++o) {
    const EnsGrain& gr = G[o];
    double u[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) u[k] = qx * gr.L[k][0] + qy * gr.L[k][1] + gr.off[k];
    cd Sflat = cd{1.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (gr.flat & (1 << k)) Sflat = Sflat * dirichlet(cd{u[k], 0.0}, sf.N[k]);
    if (gr.flat & 8) Sflat = Sflat * cexpi_d(cd{u[3], 0.0});
    cd amp = cd{0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (b >= nb) break;
      cd S = sf_branch64(gr, u, Sflat, qzb[b], sf);
      if (sf.dw != 0.0 && !sf.dw_invariant) S = exp(-sf.dw * rot_norm2(gr.rot, qx, qy, qzb[b])) * S;
      amp = amp + fcb[b] * S * cell64<SHAPE>(o, qx, qy, qzb[b], E, nelem, loc, M);
    }
    amp = gr.w * amp;
    acc += amp.r * amp.r + amp.i * amp.i;
  }
  part[(int64_t)blockIdx.y * g.npix + p] = acc;
}

// Grain-invariant cells (the host checked it: spheres and cylinders whose F depends on q only through
// q.q and the third row of R_g R_e, which every grain shares, with locations that every grain sees
// alike): the cell's form-factor sum times the branch factor, once per (pixel, branch), into ffh
// [nb][npix] complex. A kernel of its own, so that the registers and scratch of the form-factor code
// (the sphere cap's FP64 Bessel function, the FP64 fallbacks of the FP32C forms) do not size the
// structure-factor loop over the grains.
template <int SHAPE>
__global__ __launch_bounds__(256) void k_ensemble_cell64(QViewE g, int nb, const double* __restrict__ fc,
                                                         const EnsElem* __restrict__ E, int nelem,
                                                         const double* __restrict__ loc, const double* __restrict__ M,
                                                         EnsSF sf, double* __restrict__ ffh) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const double qx = g.qx[p], qy = g.qy[p];
  for (int b = 0; b < nb; ++b) {
    const int64_t q = p + b * g.npix;
    const cd qz = cd{g.qz[2 * q], g.qz[2 * q + 1]};
    const cd f = branch_factor64(fc, q, qx, qy, qz, sf) * cell64<SHAPE>(0, qx, qy, qz, E, nelem, loc, M);
    ffh[2 * q] = f.r;
    ffh[2 * q + 1] = f.i;
  }
}

// The same pre-pass for cells of one plain sphere or cylinder shape on real q (a particle in vacuum, as
// the reference's examples place it): M q is real, so the sphere takes one real sincos per node for
// (sin x - x cos x) / x^3 and one for exp(i q_z R), and the cylinder one sincos per height node for both
// sinc(q_z H / 2) and exp(i q_z H / 2), where the complex path evaluates complex sines and divisions
// (src/ff/ff_ana_sphere.cpp:33-45 and src/ff/ff_ana_cylinder.cpp:34-42 form the same kernels on complex q).
template <int SHAPE>
__device__ inline cd real_cell_ff(const double mq[3], const Nodes& P);

template <>
__device__ inline cd real_cell_ff<kShapePlainSphere>(const double mq[3], const Nodes& P) {
  const double qv = sqrt(mq[0] * mq[0] + mq[1] * mq[1] + mq[2] * mq[2]);
  cd F = cd{0.0, 0.0};
  if (qv < 2.2250738585072014e-308) return F;
  for (int i = 0; i < P.n[0]; ++i) {
    const double R = P.val[0][i], x = qv * R;
    double sx, cx, sz, cz;
    sincos(x, &sx, &cx);
    sincos(mq[2] * R, &sz, &cz);
    const double a = P.wt[0][i] * (4 * kPi * R * R * R) * ((sx - x * cx) / (x * x * x));
    F = F + cd{a * cz, a * sz};
  }
  return F;
}

template <>
__device__ inline cd real_cell_ff<HGX_SHAPE_CYLINDER>(const double mq[3], const Nodes& P) {
  const double qp = sqrt(mq[0] * mq[0] + mq[1] * mq[1]);
  cd F = cd{0.0, 0.0};
  for (int i = 0; i < P.n[0]; ++i) {
    const double R = P.val[0][i], t = qp * R;
    const double bess = t * t > 1.0e-13 ? ::j1(t) / t : 0.0;  // J1(t)/t, 0 below |t|^2 = 1e-13
    for (int k = 0; k < P.n[1]; ++k) {
      const double H = P.val[1][k], hz = 0.5 * H * mq[2];
      double sz, cz;
      sincos(hz, &sz, &cz);
      const double sinc = fabs(hz) > 1.0e-14 ? sz / hz : 1.0 - hz * hz / 6.0 + hz * hz * hz * hz / 120.0;
      const double a = (P.wt[0][i] * P.wt[1][k]) * (2. * kPi * R * R * H) * (sinc * bess);
      F = F + cd{a * cz, a * sz};
    }
  }
  return F;
}

template <int SHAPE>
__global__ __launch_bounds__(256) void k_ensemble_cell64_real(QViewE g, int nb, const double* __restrict__ fc,
                                                              const EnsElem* __restrict__ E, int nelem,
                                                              const double* __restrict__ loc,
                                                              const double* __restrict__ M, EnsSF sf,
                                                              double* __restrict__ ffh) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const double qx = g.qx[p], qy = g.qy[p];
  for (int b = 0; b < nb; ++b) {
    const int64_t q = p + b * g.npix;
    const double qz = g.qz[2 * q];
    cd cell = cd{0.0, 0.0};
    for (int e = 0; e < nelem; ++e) {
      const double* m = M + 9 * (int64_t)e;
      double mq[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) mq[j] = m[3 * j] * qx + m[3 * j + 1] * qy + m[3 * j + 2] * qz;
      const cd F = real_cell_ff<SHAPE>(mq, E[e].P);
      cd ph = cd{0.0, 0.0};
      for (int l = E[e].loc0; l < E[e].loc0 + E[e].nloc; ++l) {
        const double* t = loc + 3 * l;
        double s, c;
        sincos(mq[0] * t[0] + mq[1] * t[1] + mq[2] * t[2], &s, &c);
        ph = ph + cd{c, s};
      }
      cell = cell + cd{E[e].dn2r, E[e].dn2i} * (F * ph);
    }
    const cd f = branch_factor64(fc, q, qx, qy, cd{qz, 0.0}, sf) * cell;
    ffh[2 * q] = f.r;
    ffh[2 * q + 1] = f.i;
  }
}

// sum_b F_b S_g(q_b) for grain gr, FP64; CQ = false: real q_z, so every phase is real and the
// structure factor is a real Dirichlet product times the centre phase (real sines, one sincos).
template <int NB, bool CQ>
__device__ __attribute__((always_inline)) inline cd branch_sum64(const EnsGrain& gr, double qx, double qy, const cd (&qzb)[NB], const cd (&F)[NB],
                                  const EnsSF& sf) {
  double u[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) u[k] = qx * gr.L[k][0] + qy * gr.L[k][1] + gr.off[k];
  cd sum = cd{0.0, 0.0};
  if constexpr (CQ) {
    cd Sflat = cd{1.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (gr.flat & (1 << k)) Sflat = Sflat * dirichlet(cd{u[k], 0.0}, sf.N[k]);
    if (gr.flat & 8) Sflat = Sflat * cexpi_d(cd{u[3], 0.0});
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      cd S = sf_branch64(gr, u, Sflat, qzb[b], sf);
      if (sf.dw != 0.0 && !sf.dw_invariant) S = exp(-sf.dw * rot_norm2(gr.rot, qx, qy, qzb[b])) * S;
      sum = sum + F[b] * S;
    }
  } else {
    double Df = 1.0;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (gr.flat & (1 << k)) Df *= dirichlet_real(u[k], sf.N[k]);
    const cd Pf = (gr.flat & 8) ? expi_real(u[3]) : cd{1.0, 0.0};
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const double qz = qzb[b].r;
      double D = Df;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (!(gr.flat & (1 << k))) D *= dirichlet_real(u[k] + qz * gr.L[k][2], sf.N[k]);
      if (sf.dw != 0.0 && !sf.dw_invariant) D *= exp(-sf.dw * rot_norm2(gr.rot, qx, qy, qzb[b]));
      const cd P = (gr.flat & 8) ? Pf : expi_real(u[3] + qz * gr.L[3][2]);
      sum = sum + D * (F[b] * P);
    }
  }
  return sum;
}

// The grain loop of a grain-invariant cell: sum_g |w_g sum_b ffh_b S_g(q_b)|^2 over this y-slice.
template <int NB, bool CQ>
__global__ __launch_bounds__(256) void k_ensemble_sf64(QViewE g, int nb, const double* __restrict__ ffh,
                                                       const EnsGrain* __restrict__ G, int ngrains, int per_y,
                                                       EnsSF sf, double* __restrict__ part) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const int g0 = blockIdx.y * per_y, g1 = min(ngrains, g0 + per_y);
  const double qx = g.qx[p], qy = g.qy[p];
  cd qzb[NB], F[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int64_t q = p + b * g.npix;
    qzb[b] = cd{g.qz[2 * q], g.qz[2 * q + 1]};
    F[b] = cd{ffh[2 * q], ffh[2 * q + 1]};
  }
  double acc = 0.0;
  for (int o = g0; o < g1; ++o) {
    const EnsGrain& gr = G[o];
    const cd amp = gr.w * branch_sum64<NB, CQ>(gr, qx, qy, qzb, F, sf);
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

// FP64 phase in turns reduced to [-1/2, 1/2], as FP32
__device__ inline float red_turn(double t) { return (float)(t - rint(t)); }

// The FP32C structure factor of grain gr on branch b: the branch-independent factors Df, Pf (flat
// directions) and the rest from the FP64 phases in turns.
__device__ inline cf sf_branch32(const EnsGrain& gr, const double u[4], float Df, cf Pf, double qz, const float N[3],
                                 const EnsSF& sf, double qx, double qy) {
  float D = Df;
  if (!(gr.flat & 1)) D *= dirichlet_turn(red_turn(u[0] + qz * gr.L[0][2]), N[0]);
  if (!(gr.flat & 2)) D *= dirichlet_turn(red_turn(u[1] + qz * gr.L[1][2]), N[1]);
  if (!(gr.flat & 4)) D *= dirichlet_turn(red_turn(u[2] + qz * gr.L[2][2]), N[2]);
  const cf P = (gr.flat & 8) ? Pf : expi_turn(red_turn(u[3] + qz * gr.L[3][2]));
  if (sf.dw != 0.0 && !sf.dw_invariant) D *= __expf((float)(-sf.dw * rot_norm2(gr.rot, qx, qy, cd{qz, 0.0})));
  return D * P;
}

// the branch-independent factors of grain gr (FP32C)
__device__ inline void sf_flat32(const EnsGrain& gr, const double u[4], const float N[3], float& Df, cf& Pf) {
  Df = 1.0f;
  Pf = cf{1.0f, 0.0f};
  if (gr.flat & 1) Df *= dirichlet_turn(red_turn(u[0]), N[0]);
  if (gr.flat & 2) Df *= dirichlet_turn(red_turn(u[1]), N[1]);
  if (gr.flat & 4) Df *= dirichlet_turn(red_turn(u[2]), N[2]);
  if (gr.flat & 8) Pf = expi_turn(red_turn(u[3]));
}

__device__ inline double branch_dw32(double qx, double qy, double qz, const EnsSF& sf) {
  return (sf.dw != 0.0 && sf.dw_invariant) ? (double)__expf((float)(-sf.dw * (qx * qx + qy * qy + qz * qz))) : 1.0;
}

template <int SHAPE>
__global__ __launch_bounds__(256) void k_ensemble32(QViewE g, int nb, const double* __restrict__ fc,
                                                    const EnsElem* __restrict__ E, int nelem,
                                                    const double* __restrict__ loc, const double* __restrict__ M,
                                                    const EnsGrain* __restrict__ G, int ngrains, int per_y,
                                                    EnsSF sf, double* __restrict__ part) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const int g0 = blockIdx.y * per_y, g1 = min(ngrains, g0 + per_y);
  const double qx = g.qx[p], qy = g.qy[p];
  const float N[3] = {(float)sf.N[0], (float)sf.N[1], (float)sf.N[2]};
  double qzb[4];
  cd fcb[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if (b >= nb) break;
    const int64_t q = p + b * g.npix;
    qzb[b] = g.qz[2 * q];
    fcb[b] = branch_dw32(qx, qy, qzb[b], sf) * (fc ? cd{fc[2 * q], fc[2 * q + 1]} : cd{1.0, 0.0});
  }
  double acc = 0.0;
  for (int o = g0; o < g1; ++o) {
    const EnsGrain& gr = G[o];
    // in-plane parts of the phases, in turns (FP64)
    double u[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) u[k] = qx * gr.L[k][0] + qy * gr.L[k][1] + gr.off[k];
    float Df;
    cf Pf;
    sf_flat32(gr, u, N, Df, Pf);
    cd amp = cd{0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (b >= nb) break;
      const cf S = sf_branch32(gr, u, Df, Pf, qzb[b], N, sf, qx, qy);
      amp = amp + fcb[b] * (cd{(double)S.r, (double)S.i} * cell32<SHAPE>(o, qx, qy, qzb[b], E, nelem, loc, M));
    }
    amp = gr.w * amp;
    acc += amp.r * amp.r + amp.i * amp.i;
  }
  part[(int64_t)blockIdx.y * g.npix + p] = acc;
}

// FP32C grain-invariant cell (see k_ensemble_cell64)
template <int SHAPE>
__global__ __launch_bounds__(256) void k_ensemble_cell32(QViewE g, int nb, const double* __restrict__ fc,
                                                         const EnsElem* __restrict__ E, int nelem,
                                                         const double* __restrict__ loc, const double* __restrict__ M,
                                                         EnsSF sf, double* __restrict__ ffh) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const double qx = g.qx[p], qy = g.qy[p];
  for (int b = 0; b < nb; ++b) {
    const int64_t q = p + b * g.npix;
    const double qz = g.qz[2 * q];
    const cd f = branch_dw32(qx, qy, qz, sf) * (fc ? cd{fc[2 * q], fc[2 * q + 1]} : cd{1.0, 0.0});
    const cd v = f * cell32<SHAPE>(0, qx, qy, qz, E, nelem, loc, M);
    ffh[2 * q] = v.r;
    ffh[2 * q + 1] = v.i;
  }
}

// FP32C grain loop of a grain-invariant cell
template <int NB>
__global__ __launch_bounds__(256) void k_ensemble_sf32(QViewE g, int nb, const double* __restrict__ ffh,
                                                       const EnsGrain* __restrict__ G, int ngrains, int per_y,
                                                       EnsSF sf, double* __restrict__ part) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const int g0 = blockIdx.y * per_y, g1 = min(ngrains, g0 + per_y);
  const double qx = g.qx[p], qy = g.qy[p];
  const float N[3] = {(float)sf.N[0], (float)sf.N[1], (float)sf.N[2]};
  // the branch factors in FP32: the sum over four branches is the grain's amplitude, each term an
  // FP32 structure factor times an FP64 cell value, so FP32 products lose nothing at the FP32C bar
  double qzb[NB];
  cf F[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int64_t q = p + b * g.npix;
    qzb[b] = g.qz[2 * q];
    F[b] = cf{(float)ffh[2 * q], (float)ffh[2 * q + 1]};
  }
  double acc = 0.0;
  for (int o = g0; o < g1; ++o) {
    const EnsGrain& gr = G[o];
    double u[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) u[k] = qx * gr.L[k][0] + qy * gr.L[k][1] + gr.off[k];
    float Df;
    cf Pf;
    sf_flat32(gr, u, N, Df, Pf);
    cf amp = cf{0.0f, 0.0f};
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      amp = amp + F[b] * sf_branch32(gr, u, Df, Pf, qzb[b], N, sf, qx, qy);
    }
    const double ar = gr.w * (double)amp.r, ai = gr.w * (double)amp.i;
    acc += ar * ar + ai * ai;
  }
  part[(int64_t)blockIdx.y * g.npix + p] = acc;
}

// ----- separable cells -------------------------------------------------------------------------
//
// A box (any node distribution: its node triple sum is the product of three per-axis sums) or a
// one-node cube is F(mq) = F_par(mq_x, mq_y) F_z(mq_z). When the cell holds that one element and
// every grain's M_g = R_g R_e keeps the z axis (third row and column (0, 0, 1): rotations about z, as
// the reference's in-plane orientation ranges give), mq_z = q_z on every grain and mq_x, mq_y do not
// depend on the branch; with the element's locations at one height t_z the location sum splits the
// same way. So F_z(q_z) e^{i q_z t_z} joins the branch factor once per (pixel, branch) and the grain loop
// evaluates F_par and the in-plane location phases once per (pixel, grain) instead of the whole cell
// per (pixel, grain, branch).

template <int SHAPE>
__device__ inline double fpar64(double mx, double my, const Nodes& P) {
  if (SHAPE == HGX_SHAPE_CUBE) {
    const double L = P.val[0][0];
    return (P.wt[0][0] * L * L) * sinc_free(cd{0.5 * L * mx, 0.0}).r * sinc_free(cd{0.5 * L * my, 0.0}).r;
  }
  double a = 0.0, b = 0.0;
  for (int i = 0; i < P.n[0]; ++i) a += (P.wt[0][i] * P.val[0][i]) * sinc_free(cd{0.5 * P.val[0][i] * mx, 0.0}).r;
  for (int i = 0; i < P.n[1]; ++i) b += (P.wt[1][i] * P.val[1][i]) * sinc_free(cd{0.5 * P.val[1][i] * my, 0.0}).r;
  return a * b;
}

template <int SHAPE>
__device__ inline cd fz64(cd qz, const Nodes& P) {
  const int d = SHAPE == HGX_SHAPE_CUBE ? 0 : 2;
  const int n = SHAPE == HGX_SHAPE_CUBE ? 1 : P.n[2];
  cd acc = cd{0.0, 0.0};
  for (int k = 0; k < n; ++k) {
    const double H = P.val[d][k];
    const cd hz = (0.5 * H) * qz;
    acc = acc + ((SHAPE == HGX_SHAPE_CUBE ? 1.0 : P.wt[d][k]) * H) * (cexpi_d(hz) * sinc_free(hz));
  }
  return acc;
}

template <int SHAPE>
__device__ inline float fpar32(double mx, double my, const Nodes& P) {
  if (SHAPE == HGX_SHAPE_CUBE) {
    const double L = P.val[0][0];
    return (float)(P.wt[0][0] * L * L) * sinc32(mx, L) * sinc32(my, L);
  }
  float a = 0.0f, b = 0.0f;
  for (int i = 0; i < P.n[0]; ++i) a += (float)(P.wt[0][i] * P.val[0][i]) * sinc32(mx, P.val[0][i]);
  for (int i = 0; i < P.n[1]; ++i) b += (float)(P.wt[1][i] * P.val[1][i]) * sinc32(my, P.val[1][i]);
  return a * b;
}

template <int SHAPE>
__device__ inline cf fz32(double qz, const Nodes& P) {
  const int d = SHAPE == HGX_SHAPE_CUBE ? 0 : 2;
  const int n = SHAPE == HGX_SHAPE_CUBE ? 1 : P.n[2];
  cf acc = cf{0.0f, 0.0f};
  for (int k = 0; k < n; ++k) {
    const double H = P.val[d][k];
    acc = acc + ((float)((SHAPE == HGX_SHAPE_CUBE ? 1.0 : P.wt[d][k]) * H) * sinc32(qz, H)) *
                    expi_turn(turns_of(0.5 * H * qz));
  }
  return acc;
}

// Branch factors of a separable cell: fc_b, the Debye-Waller factor of orthogonal grains, dn2,
// F_z(q_z) and e^{i q_z t_z}, into ffh [nb][npix] complex.
template <int SHAPE>
__global__ __launch_bounds__(256) void k_ensemble_zcell64(QViewE g, int nb, const double* __restrict__ fc,
                                                          const EnsElem* __restrict__ E, double tz, EnsSF sf,
                                                          double* __restrict__ ffh) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const double qx = g.qx[p], qy = g.qy[p];
  const cd dn2 = cd{E[0].dn2r, E[0].dn2i};
  for (int b = 0; b < nb; ++b) {
    const int64_t q = p + b * g.npix;
    const cd qz = cd{g.qz[2 * q], g.qz[2 * q + 1]};
    const cd v = branch_factor64(fc, q, qx, qy, qz, sf) * dn2 * fz64<SHAPE>(qz, E[0].P) * cexpi_d(tz * qz);
    ffh[2 * q] = v.r;
    ffh[2 * q + 1] = v.i;
  }
}

template <int SHAPE>
__global__ __launch_bounds__(256) void k_ensemble_zcell32(QViewE g, int nb, const double* __restrict__ fc,
                                                          const EnsElem* __restrict__ E, double tz, EnsSF sf,
                                                          double* __restrict__ ffh) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const double qx = g.qx[p], qy = g.qy[p];
  const cd dn2 = cd{E[0].dn2r, E[0].dn2i};
  for (int b = 0; b < nb; ++b) {
    const int64_t q = p + b * g.npix;
    const double qz = g.qz[2 * q];
    const cd f = branch_dw32(qx, qy, qz, sf) * (fc ? cd{fc[2 * q], fc[2 * q + 1]} : cd{1.0, 0.0});
    const cf z = fz32<SHAPE>(qz, E[0].P) * expi_turn(turns_of(tz * qz));
    const cd v = f * dn2 * cd{(double)z.r, (double)z.i};
    ffh[2 * q] = v.r;
    ffh[2 * q + 1] = v.i;
  }
}

// The grain loop of a separable cell, FP64: sum_g |w_g F_par(g) L(g) sum_b ffh_b S_g(q_b)|^2 with
// L(g) the in-plane location phases.
template <int SHAPE, int NB, bool CQ>
__global__ __launch_bounds__(256) void k_ensemble_sep64(QViewE g, int nb, const double* __restrict__ ffh,
                                                        const EnsElem* __restrict__ E, const double* __restrict__ loc,
                                                        const double* __restrict__ M, const EnsGrain* __restrict__ G,
                                                        int ngrains, int per_y, EnsSF sf, double* __restrict__ part) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const int g0 = blockIdx.y * per_y, g1 = min(ngrains, g0 + per_y);
  const double qx = g.qx[p], qy = g.qy[p];
  cd qzb[NB], F[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int64_t q = p + b * g.npix;
    qzb[b] = cd{g.qz[2 * q], g.qz[2 * q + 1]};
    F[b] = cd{ffh[2 * q], ffh[2 * q + 1]};
  }
  const int l0 = E[0].loc0, l1 = E[0].loc0 + E[0].nloc;
  double acc = 0.0;
  for (int o = g0; o < g1; ++o) {
    const EnsGrain& gr = G[o];
    const double* Mo = M + 9 * (int64_t)o;
    const double mx = Mo[0] * qx + Mo[1] * qy, my = Mo[3] * qx + Mo[4] * qy;
    cd lph = cd{0.0, 0.0};
    for (int l = l0; l < l1; ++l) lph = lph + cexpi_d(cd{mx * loc[3 * l] + my * loc[3 * l + 1], 0.0});
    const cd A = fpar64<SHAPE>(mx, my, E[0].P) * lph;
    const cd sum = branch_sum64<NB, CQ>(gr, qx, qy, qzb, F, sf);
    const cd amp = gr.w * (A * sum);
    acc += amp.r * amp.r + amp.i * amp.i;
  }
  part[(int64_t)blockIdx.y * g.npix + p] = acc;
}

template <int SHAPE, int NB>
__global__ __launch_bounds__(256) void k_ensemble_sep32(QViewE g, int nb, const double* __restrict__ ffh,
                                                        const EnsElem* __restrict__ E, const double* __restrict__ loc,
                                                        const double* __restrict__ M, const EnsGrain* __restrict__ G,
                                                        int ngrains, int per_y, EnsSF sf, double* __restrict__ part) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const int g0 = blockIdx.y * per_y, g1 = min(ngrains, g0 + per_y);
  const double qx = g.qx[p], qy = g.qy[p];
  const float N[3] = {(float)sf.N[0], (float)sf.N[1], (float)sf.N[2]};
  double qzb[NB];
  cf F[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int64_t q = p + b * g.npix;
    qzb[b] = g.qz[2 * q];
    F[b] = cf{(float)ffh[2 * q], (float)ffh[2 * q + 1]};
  }
  const int l0 = E[0].loc0, l1 = E[0].loc0 + E[0].nloc;
  double acc = 0.0;
  for (int o = g0; o < g1; ++o) {
    const EnsGrain& gr = G[o];
    const double* Mo = M + 9 * (int64_t)o;
    const double mx = Mo[0] * qx + Mo[1] * qy, my = Mo[3] * qx + Mo[4] * qy;
    cf lph = cf{0.0f, 0.0f};
    for (int l = l0; l < l1; ++l) lph = lph + expi_turn(turns_of(mx * loc[3 * l] + my * loc[3 * l + 1]));
    const float fp = fpar32<SHAPE>(mx, my, E[0].P);
    double u[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) u[k] = qx * gr.L[k][0] + qy * gr.L[k][1] + gr.off[k];
    float Df;
    cf Pf;
    sf_flat32(gr, u, N, Df, Pf);
    cf sum = cf{0.0f, 0.0f};
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      sum = sum + F[b] * sf_branch32(gr, u, Df, Pf, qzb[b], N, sf, qx, qy);
    }
    const cf a = (fp * lph) * sum;
    const double ar = gr.w * (double)a.r, ai = gr.w * (double)a.i;
    acc += ar * ar + ai * ai;
  }
  part[(int64_t)blockIdx.y * g.npix + p] = acc;
}

// ----- forward-mode tangents (FP64) -------------------------------------------------------------

// One element with dual nodes, contrast and its location range; grains with dual rotation, centre
// and weight. Every table of the model enters as value + tangent along one direction.
struct EnsElemD {
  NodesT<dd> P;
  int shape;
  int loc0, nloc;
  dd dn2r, dn2i;
};
struct EnsGrainD {
  dd rot[9];
  dd center[3];
  dd w;
};

__device__ inline dz ff_any_d(int shape, const dz mq[3], const NodesT<dd>& P) {
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

// mq = M q for a dual matrix and a fixed q-point
__device__ inline void rotate_md(const dd* __restrict__ m, double qx, double qy, cd qz, dz mq[3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const dd a = m[3 * j], b = m[3 * j + 1], c = m[3 * j + 2];
    mq[j].v = cd{a.v * qx + b.v * qy + c.v * qz.r, c.v * qz.i};
    mq[j].d = cd{a.d * qx + b.d * qy + c.d * qz.r, c.d * qz.i};
  }
}

// I[p] and dI[j][p] = d/dt I[p](tables + t tangents_j) for direction j = blockIdx.y: the grain, element,
// location, structure-factor and branch sums of k_ensemble64 in dual arithmetic, every grain in one
// thread, dI = sum_g 2 Re(conj(A_g) dA_g) with A_g = w_g sum_b fc_b S_g F_g.
template <int SHAPE>
__global__ __launch_bounds__(256) void k_ensemble_tangent(QViewE g, int nb, const double* __restrict__ fc,
                                                          const EnsElemD* __restrict__ E, int nelem,
                                                          const dd* __restrict__ loc, int nloc_all,
                                                          const dd* __restrict__ M, const EnsGrainD* __restrict__ G,
                                                          int ngrains, const SFArgsT<dd>* __restrict__ SF,
                                                          double* __restrict__ I, double* __restrict__ dI) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npix) return;
  const int j = blockIdx.y;
  const EnsElemD* Ej = E + (int64_t)j * nelem;
  const dd* locj = loc + (int64_t)j * 3 * nloc_all;
  const dd* Mj = M + (int64_t)j * 9 * ngrains * nelem;
  const EnsGrainD* Gj = G + (int64_t)j * ngrains;
  const double qx = g.qx[p], qy = g.qy[p];
  double acc = 0.0, dacc = 0.0;
  for (int o = 0; o < ngrains; ++o) {
    SFArgsT<dd> sf = SF[j];
    sf.center[0] = Gj[o].center[0];
    sf.center[1] = Gj[o].center[1];
    sf.center[2] = Gj[o].center[2];
    dz amp = cconst<dz>(0.0);
    for (int b = 0; b < nb; ++b) {
      const int64_t q = p + b * g.npix;
      const cd qz = cd{g.qz[2 * q], g.qz[2 * q + 1]};
      const cd fcb = fc ? cd{fc[2 * q], fc[2 * q + 1]} : cd{1.0, 0.0};
      dz mg[3];
      rotate_md(Gj[o].rot, qx, qy, qz, mg);
      const dz S = sf_value(mg, sf);
      dz ff = cconst<dz>(0.0);
      for (int e = 0; e < nelem; ++e) {
        dz mq[3];
        rotate_md(Mj + 9 * ((int64_t)o * nelem + e), qx, qy, qz, mq);
        const dz F = SHAPE >= 0 ? shape_ff<(SHAPE >= 0 ? SHAPE : 0)>(mq, Ej[e].P) : ff_any_d(Ej[e].shape, mq, Ej[e].P);
        const dz dn2 = mkc(Ej[e].dn2r, Ej[e].dn2i);
        for (int l = Ej[e].loc0; l < Ej[e].loc0 + Ej[e].nloc; ++l) {
          const dd* t = locj + 3 * l;
          ff = ff + dn2 * (F * cexpi_d(mq[0] * t[0] + mq[1] * t[1] + mq[2] * t[2]));
        }
      }
      amp = amp + fcb * (S * ff);
    }
    amp = Gj[o].w * amp;
    acc += norm2(amp.v);
    dacc += 2.0 * (amp.v.r * amp.d.r + amp.v.i * amp.d.i);
  }
  if (j == 0) I[p] = acc;
  dI[(int64_t)j * g.npix + p] = dacc;
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

// Whether the cell is one box (or one-node cube) element whose F(M_g q) = F_par(mq_x, mq_y) F_z(q_z) on
// every grain (each M_g keeps the z axis to 1e-15) with all its locations at one height (*tz).
static bool cell_is_separable(const hgx_ensemble_model* m, const std::vector<double>& Mm, double* tz) {
  if (m->nelem != 1) return false;
  const hgx_ensemble_element& e = m->elem[0];
  if (!(e.shape == HGX_SHAPE_BOX || (e.shape == HGX_SHAPE_CUBE && e.params.n[0] == 1))) return false;
  if (e.nloc < 1 || !e.loc) return false;
  for (int l = 1; l < e.nloc; ++l)
    if (e.loc[3 * l + 2] != e.loc[2]) return false;
  for (int o = 0; o < m->ngrains; ++o) {
    const double* M = &Mm[9 * (size_t)o];
    if (std::fabs(M[2]) > 1e-15 || std::fabs(M[5]) > 1e-15 || std::fabs(M[6]) > 1e-15 || std::fabs(M[7]) > 1e-15 ||
        std::fabs(M[8] - 1.0) > 1e-15)
      return false;
  }
  *tz = e.loc[2];
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
      if (d >= analytic_dims(el.shape, el.params.flags)) continue;
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
    if (!analytic_heights_ok(el.shape, el.params.flags, el.params.n, el.params.val))
      return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_intensity: element %d: truncated sphere heights must lie in [0, 2 R]", e);
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
  // spheres without a cap run the lean instantiation (no truncated-sphere code)
  bool capped = false;
  for (int e = 0; e < ne; ++e)
    capped = capped || (E[e].shape == HGX_SHAPE_SPHERE && (E[e].P.flags & HGX_SHAPE_TRUNCATED_SPHERE) && E[e].P.n[1] > 0);
  if (common == HGX_SHAPE_SPHERE && !capped) common = kShapePlainSphere;
  if (precision == HGX_PREC_FP32C && !c->qz_real)
    return set_error(c, HGX_ERR_UNSUPPORTED, "ensemble_intensity: FP32C needs real q");
  // M[g][e] = R_g R_e with the summation order of the reference's RotMatrix_t product
  std::vector<double> Mm(9 * (size_t)ng * ne);
  std::vector<EnsGrain> G(ng);
  // the grains' phase tables: R_g^T l_k, R_g^T c_g (turns for FP32C), the compat constant, flat bits
  const double scale = precision == HGX_PREC_FP32C ? 1.0 / (2.0 * kPiD) : 1.0;
  const bool compat_lb2 = (m->sf_flags & HGX_SF_COMPAT_LB2) != 0;
  bool orthogonal = true;
  for (int o = 0; o < ng; ++o) {
    const double* Rg = m->rot + 9 * (size_t)o;
    EnsGrain& gr = G[o];
    std::memset(&gr, 0, sizeof(gr));
    for (int k = 0; k < 9; ++k) gr.rot[k] = Rg[k];
    double vec[4][3];
    for (int k = 0; k < 3; ++k) {
      vec[0][k] = m->la[k];
      vec[1][k] = m->lb[k];
      vec[2][k] = m->lc[k];
      vec[3][k] = m->center ? m->center[3 * (size_t)o + k] : 0.0;
    }
    if (compat_lb2) {
      // src/sf/sf.cpp:212: lb[2] * lb[2] in place of lb[2] * mq[2], a constant phase
      gr.off[1] = m->lb[2] * m->lb[2] * scale;
      vec[1][2] = 0.0;
    }
    for (int v = 0; v < 4; ++v)
      for (int k = 0; k < 3; ++k)  // (R^T v)_k = sum_j R[j][k] v_j
        gr.L[v][k] = (Rg[k] * vec[v][0] + Rg[3 + k] * vec[v][1] + Rg[6 + k] * vec[v][2]) * scale;
    for (int v = 0; v < 4; ++v)
      if (gr.L[v][2] == 0.0) gr.flat |= 1 << v;
    gr.w = m->weight ? m->weight[o] : 1.0;
    for (int i = 0; i < 3 && orthogonal; ++i)
      for (int j = 0; j < 3; ++j) {
        double d = 0.0;
        for (int k = 0; k < 3; ++k) d += Rg[3 * k + i] * Rg[3 * k + j];
        if (std::fabs(d - (i == j ? 1.0 : 0.0)) > 1e-12) orthogonal = false;
      }
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
  double tz = 0.0;
  const bool sep = !hoist && cell_is_separable(m, Mm, &tz);
  EnsSF sfb;
  for (int k = 0; k < 3; ++k) sfb.N[k] = m->repet[k] < 1 ? 1 : m->repet[k];  // src/sf/sf.cpp:165-167
  sfb.dw = 0.5 * m->dw_sigma * m->dw_sigma;
  sfb.dw_invariant = orthogonal ? 1 : 0;
  if (int rc = set_device(c)) return rc;
  // grains split over blockIdx.y so that a launch has at least ~2048 workgroups
  const unsigned nbx = ceil_div(c->npix, 256);
  int ny = std::max(1, std::min(ng, (int)((2048 + nbx - 1) / nbx)));
  const int per_y = (ng + ny - 1) / ny;
  ny = (ng + per_y - 1) / per_y;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t b_part = al((size_t)ny * c->npix * sizeof(double)), b_E = al(ne * sizeof(EnsElem)),
               b_loc = al(loc.size() * sizeof(double)), b_M = al(Mm.size() * sizeof(double)),
               b_G = al(ng * sizeof(EnsGrain)), b_ffh = hoist || sep ? al((size_t)c->nq * 2 * sizeof(double)) : 0;
  if (int rc = ensure_work(c, b_part + b_E + b_loc + b_M + b_G + b_ffh)) return rc;
  char* w = static_cast<char*>(c->d_work);
  double* part = reinterpret_cast<double*>(w);
  EnsElem* dE = reinterpret_cast<EnsElem*>(w + b_part);
  double* dloc = reinterpret_cast<double*>(w + b_part + b_E);
  double* dM = reinterpret_cast<double*>(w + b_part + b_E + b_loc);
  EnsGrain* dG = reinterpret_cast<EnsGrain*>(w + b_part + b_E + b_loc + b_M);
  double* ffh = hoist || sep ? reinterpret_cast<double*>(w + b_part + b_E + b_loc + b_M + b_G) : nullptr;
  {
    const void* src[4] = {E.data(), loc.data(), Mm.data(), G.data()};
    const size_t nb[4] = {ne * sizeof(EnsElem), loc.size() * sizeof(double), Mm.size() * sizeof(double),
                          ng * sizeof(EnsGrain)};
    if (int rc = stage_to_device(c, 4, src, nb, reinterpret_cast<char*>(dE))) return rc;
  }
  QViewE g{c->d_qx, c->d_qy, c->d_qz, c->npix};
  const dim3 grid(nbx, ny);
  const bool f32 = precision == HGX_PREC_FP32C;
  if (hoist) {
    {
      LaunchTimer lt(c, true);
#define HGX_ENS_CELL(S)                                                                                     \
  if (f32) k_ensemble_cell32<S><<<nbx, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, ne, dloc, dM, sfb, ffh); \
  else k_ensemble_cell64<S><<<nbx, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, ne, dloc, dM, sfb, ffh)
      switch (common) {
        case kShapePlainSphere:
          if (!f32 && c->qz_real)
            k_ensemble_cell64_real<kShapePlainSphere>
                <<<nbx, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, ne, dloc, dM, sfb, ffh);
          else HGX_ENS_CELL(kShapePlainSphere);
          break;
        case HGX_SHAPE_SPHERE: HGX_ENS_CELL(HGX_SHAPE_SPHERE); break;
        case HGX_SHAPE_CYLINDER:
          if (!f32 && c->qz_real)
            k_ensemble_cell64_real<HGX_SHAPE_CYLINDER>
                <<<nbx, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, ne, dloc, dM, sfb, ffh);
          else HGX_ENS_CELL(HGX_SHAPE_CYLINDER);
          break;
        default: HGX_ENS_CELL(-1); break;
      }
#undef HGX_ENS_CELL
      HGX_CHECK_LAUNCH(c);
    }
    LaunchTimer lt(c, true);
    const bool four = c->nbranch == 4;
    if (f32 && four) k_ensemble_sf32<4><<<grid, 256, 0, c->stream>>>(g, 4, ffh, dG, ng, per_y, sfb, part);
    else if (f32) k_ensemble_sf32<1><<<grid, 256, 0, c->stream>>>(g, 1, ffh, dG, ng, per_y, sfb, part);
    else if (four && c->qz_real) k_ensemble_sf64<4, false><<<grid, 256, 0, c->stream>>>(g, 4, ffh, dG, ng, per_y, sfb, part);
    else if (four) k_ensemble_sf64<4, true><<<grid, 256, 0, c->stream>>>(g, 4, ffh, dG, ng, per_y, sfb, part);
    else if (c->qz_real) k_ensemble_sf64<1, false><<<grid, 256, 0, c->stream>>>(g, 1, ffh, dG, ng, per_y, sfb, part);
    else k_ensemble_sf64<1, true><<<grid, 256, 0, c->stream>>>(g, 1, ffh, dG, ng, per_y, sfb, part);
    HGX_CHECK_LAUNCH(c);
  } else if (sep) {
    const int sh = m->elem[0].shape;
    {
      LaunchTimer lt(c, true);
#define HGX_ENS_Z(S)                                                                                   \
  if (f32) k_ensemble_zcell32<S><<<nbx, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, tz, sfb, ffh); \
  else k_ensemble_zcell64<S><<<nbx, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, tz, sfb, ffh)
      if (sh == HGX_SHAPE_BOX) HGX_ENS_Z(HGX_SHAPE_BOX);
      else HGX_ENS_Z(HGX_SHAPE_CUBE);
#undef HGX_ENS_Z
      HGX_CHECK_LAUNCH(c);
    }
    LaunchTimer lt(c, true);
#define HGX_ENS_SEP_NB(S, B)                                                                                    \
  if (f32) k_ensemble_sep32<S, B><<<grid, 256, 0, c->stream>>>(g, B, ffh, dE, dloc, dM, dG, ng, per_y, sfb, part); \
  else if (c->qz_real) k_ensemble_sep64<S, B, false><<<grid, 256, 0, c->stream>>>(g, B, ffh, dE, dloc, dM, dG, ng, per_y, sfb, part); \
  else k_ensemble_sep64<S, B, true><<<grid, 256, 0, c->stream>>>(g, B, ffh, dE, dloc, dM, dG, ng, per_y, sfb, part)
#define HGX_ENS_SEP(S)                  \
  if (c->nbranch == 4) {                \
    HGX_ENS_SEP_NB(S, 4);               \
  } else {                              \
    HGX_ENS_SEP_NB(S, 1);               \
  }
    if (sh == HGX_SHAPE_BOX) {
      HGX_ENS_SEP(HGX_SHAPE_BOX)
    } else {
      HGX_ENS_SEP(HGX_SHAPE_CUBE)
    }
#undef HGX_ENS_SEP
#undef HGX_ENS_SEP_NB
    HGX_CHECK_LAUNCH(c);
  } else {
    LaunchTimer lt(c, true);
#define HGX_ENS_LAUNCH(S)                                                                                       \
  if (f32) k_ensemble32<S><<<grid, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, ne, dloc, dM, dG, ng, per_y, sfb, part); \
  else k_ensemble64<S><<<grid, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, ne, dloc, dM, dG, ng, per_y, sfb, part)
    switch (common) {
      case kShapePlainSphere: HGX_ENS_LAUNCH(kShapePlainSphere); break;
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

// The tangent pass: the value tables as ensemble_run builds them, each direction's tangents beside
// them (the same counts), dual tables staged in one go, one launch over (pixels, directions).
static int ensemble_tangent_run(Context* c, const hgx_ensemble_model* m, int ndir, const hgx_ensemble_model* t,
                                const double* fc, double* I, double* dI) {
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "ensemble_tangent: no q-grid");
  if (!m || !t || !I || !dI || ndir < 1 || ndir > 64 || m->nelem < 1 || m->nelem > 64 || !m->elem ||
      m->ngrains < 1 || !m->rot)
    return set_error(c, HGX_ERR_INVALID_ARG,
                     "ensemble_tangent: need a model (1..64 elements, >= 1 grain), 1..64 tangents and outputs");
  const int ne = m->nelem, ng = m->ngrains;
  if ((int64_t)ng * ne * ndir > ((int64_t)1 << 22))
    return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_tangent: grains x elements x directions above 2^22");
  int nloc_all = 0;
  for (int e = 0; e < ne; ++e) {
    const hgx_ensemble_element& el = m->elem[e];
    if (el.shape < 0 || el.shape > HGX_SHAPE_PYRAMID || el.nloc < 1 || !el.loc)
      return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_tangent: element %d needs a shape and locations", e);
    if (el.shape == HGX_SHAPE_SPHERE && (el.params.flags & HGX_SHAPE_TRUNCATED_SPHERE))
      return set_error(c, HGX_ERR_UNSUPPORTED, "ensemble_tangent: element %d: truncated spheres have no tangent", e);
    for (int d = 0; d < analytic_dims(el.shape, el.params.flags); ++d)
      if (el.params.n[d] < 1 || el.params.n[d] > HGX_MAX_NODES || !el.params.val[d] || !el.params.wt[d])
        return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_tangent: element %d dimension %d needs 1..%d nodes", e, d,
                         HGX_MAX_NODES);
    nloc_all += el.nloc;
  }
  // every tangent model has the value model's counts; node and location arrays may be NULL (zero)
  for (int j = 0; j < ndir; ++j) {
    const hgx_ensemble_model& tj = t[j];
    if (tj.nelem != ne || tj.ngrains != ng || !tj.elem)
      return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_tangent: tangent %d has other counts than the model", j);
    for (int e = 0; e < ne; ++e) {
      if (tj.elem[e].shape != m->elem[e].shape || tj.elem[e].nloc != m->elem[e].nloc)
        return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_tangent: tangent %d element %d differs in shape or locations",
                         j, e);
      for (int d = 0; d < analytic_dims(m->elem[e].shape, m->elem[e].params.flags); ++d)
        if ((tj.elem[e].params.val[d] || tj.elem[e].params.wt[d]) && tj.elem[e].params.n[d] != m->elem[e].params.n[d])
          return set_error(c, HGX_ERR_INVALID_ARG, "ensemble_tangent: tangent %d element %d dimension %d node count", j,
                           e, d);
    }
  }
  std::vector<EnsElemD> E((size_t)ndir * ne);
  std::vector<dd> loc((size_t)ndir * 3 * nloc_all), M((size_t)ndir * 9 * ng * ne);
  std::vector<EnsGrainD> G((size_t)ndir * ng);
  std::vector<SFArgsT<dd>> SF(ndir);
  int common = m->elem[0].shape;
  for (int e = 1; e < ne; ++e)
    if (m->elem[e].shape != common) common = -1;
  for (int j = 0; j < ndir; ++j) {
    const hgx_ensemble_model& tj = t[j];
    int l0 = 0;
    for (int e = 0; e < ne; ++e) {
      const hgx_ensemble_element &el = m->elem[e], &te = tj.elem[e];
      EnsElemD& D = E[(size_t)j * ne + e];
      std::memset(&D, 0, sizeof(D));
      for (int d = 0; d < 4; ++d) {
        if (d >= analytic_dims(el.shape, el.params.flags)) continue;
        D.P.n[d] = el.params.n[d];
        for (int i = 0; i < el.params.n[d]; ++i) {
          D.P.val[d][i] = dd{el.params.val[d][i], te.params.val[d] ? te.params.val[d][i] : 0.0};
          D.P.wt[d][i] = dd{el.params.wt[d][i], te.params.wt[d] ? te.params.wt[d][i] : 0.0};
        }
      }
      const double tt = std::tan(el.params.tau);
      D.P.tan_tau = dd{tt, (1.0 + tt * tt) * te.params.tau};
      D.P.sin_eta = dd{std::sin(el.params.eta), std::cos(el.params.eta) * te.params.eta};
      D.P.cos_eta = dd{std::cos(el.params.eta), -std::sin(el.params.eta) * te.params.eta};
      D.P.flags = el.params.flags;
      D.shape = el.shape;
      D.loc0 = l0;
      D.nloc = el.nloc;
      D.dn2r = dd{el.dn2_re, te.dn2_re};
      D.dn2i = dd{el.dn2_im, te.dn2_im};
      for (int k = 0; k < 3 * el.nloc; ++k)
        loc[(size_t)j * 3 * nloc_all + 3 * (size_t)l0 + k] = dd{el.loc[k], te.loc ? te.loc[k] : 0.0};
      l0 += el.nloc;
    }
    for (int o = 0; o < ng; ++o) {
      const double* Rg = m->rot + 9 * (size_t)o;
      const double* dRg = tj.rot ? tj.rot + 9 * (size_t)o : nullptr;
      EnsGrainD& Gd = G[(size_t)j * ng + o];
      for (int k = 0; k < 9; ++k) Gd.rot[k] = dd{Rg[k], dRg ? dRg[k] : 0.0};
      for (int k = 0; k < 3; ++k)
        Gd.center[k] = dd{m->center ? m->center[3 * (size_t)o + k] : 0.0, tj.center ? tj.center[3 * (size_t)o + k] : 0.0};
      Gd.w = dd{m->weight ? m->weight[o] : 1.0, tj.weight ? tj.weight[o] : 0.0};
      for (int e = 0; e < ne; ++e) {
        // M = R_g R_e (the order of ensemble_run), dM = dR_g R_e + R_g dR_e
        const double* Re = m->elem[e].rot;
        const double* dRe = tj.elem[e].rot;
        dd* out = &M[(size_t)j * 9 * ng * ne + 9 * ((size_t)o * ne + e)];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            double v = 0, dv = 0;
            for (int k = 0; k < 3; ++k) {
              v += Rg[3 * a + k] * Re[3 * k + b];
              dv += (dRg ? dRg[3 * a + k] : 0.0) * Re[3 * k + b] + Rg[3 * a + k] * dRe[3 * k + b];
            }
            out[3 * a + b] = dd{v, dv};
          }
      }
    }
    SFArgsT<dd>& sf = SF[j];
    for (int k = 0; k < 3; ++k) {
      sf.la[k] = dd{m->la[k], tj.la[k]};
      sf.lb[k] = dd{m->lb[k], tj.lb[k]};
      sf.lc[k] = dd{m->lc[k], tj.lc[k]};
      sf.N[k] = m->repet[k] < 1 ? 1 : m->repet[k];
      sf.center[k] = dd{0.0, 0.0};
    }
    sf.dw = dd{0.5 * m->dw_sigma * m->dw_sigma, m->dw_sigma * tj.dw_sigma};
    sf.compat_lb2 = (m->sf_flags & HGX_SF_COMPAT_LB2) != 0;
  }
  if (int rc = set_device(c)) return rc;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t b_E = al(E.size() * sizeof(EnsElemD)), b_loc = al(loc.size() * sizeof(dd)),
               b_M = al(M.size() * sizeof(dd)), b_G = al(G.size() * sizeof(EnsGrainD)),
               b_SF = al(SF.size() * sizeof(SFArgsT<dd>));
  if (int rc = ensure_work(c, b_E + b_loc + b_M + b_G + b_SF)) return rc;
  char* w = static_cast<char*>(c->d_work);
  EnsElemD* dE = reinterpret_cast<EnsElemD*>(w);
  dd* dloc = reinterpret_cast<dd*>(w + b_E);
  dd* dM = reinterpret_cast<dd*>(w + b_E + b_loc);
  EnsGrainD* dG = reinterpret_cast<EnsGrainD*>(w + b_E + b_loc + b_M);
  SFArgsT<dd>* dSF = reinterpret_cast<SFArgsT<dd>*>(w + b_E + b_loc + b_M + b_G);
  {
    const void* src[5] = {E.data(), loc.data(), M.data(), G.data(), SF.data()};
    const size_t nb[5] = {E.size() * sizeof(EnsElemD), loc.size() * sizeof(dd), M.size() * sizeof(dd),
                          G.size() * sizeof(EnsGrainD), SF.size() * sizeof(SFArgsT<dd>)};
    if (int rc = stage_to_device(c, 5, src, nb, w)) return rc;
  }
  QViewE g{c->d_qx, c->d_qy, c->d_qz, c->npix};
  LaunchTimer lt(c, true);
  const dim3 grid(ceil_div(c->npix, 256), ndir);
#define HGX_TAN_LAUNCH(S) \
  k_ensemble_tangent<S><<<grid, 256, 0, c->stream>>>(g, c->nbranch, fc, dE, ne, dloc, nloc_all, dM, dG, ng, dSF, I, dI)
  switch (common) {
    case HGX_SHAPE_SPHERE: HGX_TAN_LAUNCH(HGX_SHAPE_SPHERE); break;
    case HGX_SHAPE_CYLINDER: HGX_TAN_LAUNCH(HGX_SHAPE_CYLINDER); break;
    case HGX_SHAPE_BOX: HGX_TAN_LAUNCH(HGX_SHAPE_BOX); break;
    default: HGX_TAN_LAUNCH(-1); break;
  }
#undef HGX_TAN_LAUNCH
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

}  // namespace hgx

extern "C" int hgx_ensemble_intensity_tangent(hgx_ctx* h, const hgx_ensemble_model* model, int ndir,
                                              const hgx_ensemble_model* tangents, const double* fc, double* I,
                                              double* dI) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  return ensemble_tangent_run(c, model, ndir, tangents, fc, I, dI);
}

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
