// grating.hip -- batched fit of trapezoidal grating lines: device meshing, FP32C numeric form factor
// on the matrix cores, chi^2, its gradient through the record adjoint, and L-BFGS, all on the device.
//
// Model (theta = top width wt, sidewall angle alpha (rad), height h, scale s): a line of trapezoidal
// cross-section in the y-z plane (base on z = 0, bottom width wb = wt + 2 h cot(alpha)) extruded along
// x over `length`, repeated `periods` times along y with `pitch`:
//     I(p) = s SF(qy)^2 |dn2 sum_b fc_b(p) F_b(p)|^2,  SF(qy) = sin(N qy d / 2) / sin(qy d / 2)
//     chi^2 = sum_p m_p (I_p - O_p)^2 / max(O_p, 1)
// F is the approximate-triangle form factor (src/ff/cpu/ff_tri_cpu.cpp:200-246) of the line's mesh, the
// branch sum the DWBA combination of src/sim/hipgisaxs_main.cpp:1084-1100 and SF the centred form of the
// lattice sum of src/sf/sf.cpp:203-232 along one axis. The reference fits such models with TAO/PSO and
// finite-difference gradients (src/analyzer/hipgisaxs_fit_lmvm.cpp:190-224); here every candidate's
// gradient comes from one adjoint pass:
//     dchi^2/dtheta_j = sum_t dL/dA_t . dA_t/dtheta_j + dL/dc_t . dc_t/dtheta_j,
// with L = Re sum_q W F, W_b = 4 s SF^2 m (I - O) / max(O, 1) conj(A) dn2 fc_b (A the amplitude) and
// (dL/dA_t, dL/dc_t) from the product-grid adjoint kernel (adjoint.hip). dA_t/dtheta and dc_t/dtheta
// are exact: every vertex is a fixed convex combination of the cross-section corners.
//
// Mesh: perimeter of n_bottom + 2 n_side + n_top segments (a multiple of 16), n_length segments along
// x (a multiple of 8), two triangles per quad, caps fanned from the corner centroid. Slots are laid
// out in 128-triangle blocks of 8 x 8 quads (spatially compact for the FP32 in-block phases) and cap
// blocks of 16 fan triangles plus zero padding (A = 0: contributes exactly nothing).
//
// Candidates run in groups of G that share device buffers; one group's step is mesh -> block origins
// -> forward -> pixel pass (intensity, chi^2, adjoint weights) -> adjoint -> gradient contraction.
// The whole evaluation (all groups) and the L-BFGS update are captured once into a hipGraph.
#include <algorithm>
#include <cmath>

#include "adjoint_core.hpp"
#include "numeric_core.hpp"

namespace hgx {

constexpr int kLbfgsMem = 6;
constexpr int kGPar = 4;          // wt, alpha, h, scale
constexpr int kGChunkBlocks = 64; // forward chunk (FP32 sums of at most 64 blocks)

struct GMesh {
  double L;
  int nb, ns, nt, nx, np;
  int nlb, ncb, nblocks;
  int64_t nslots;
};

// Per-candidate parameters and optimiser state (device).
struct GState {
  double th[kGPar];     // accepted point
  double f, g[kGPar];   // chi^2 and gradient there
  double trial[kGPar];  // point being evaluated
  double ft, gt[kGPar]; // chi^2 and gradient at the trial point
  double d[kGPar];      // search direction (scaled coordinates)
  double step;
  double ref[kGPar];    // parameter scales: u = theta / ref
  double S[kLbfgsMem][kGPar], Y[kLbfgsMem][kGPar], rho[kLbfgsMem];
  int nhist, head, iter, status;
};

struct GPixArgs {
  const double* qx;
  const double* qy;
  const double* qz;
  int64_t npix;
  int nbranch;
  const double* fc;    // [2 * nbranch * npix] or NULL (Born, one branch)
  const double* obs;   // [npix]
  const double* mask;  // [npix] or NULL
  double dn2r, dn2i;
  double pitch;
  int periods;
  double L;
};

// ----- geometry -------------------------------------------------------------------------------

// Perimeter vertex k (counter-clockwise in the y-z plane: bottom C0 -> C1, right side C1 -> C2, top
// C2 -> C3, left side C3 -> C0) and its derivatives with respect to (wt, alpha, h).
__device__ inline void g_perim(const GMesh& gm, int k, const double th[3], double P[2], double dP[3][2]) {
  const double wt = th[0], al = th[1], h = th[2];
  double sa, ca;
  sincos(al, &sa, &ca);
  const double ct = ca / sa, cs2 = 1.0 / (sa * sa);
  const double wb = wt + 2.0 * h * ct;
  const double C[4][2] = {{-0.5 * wb, 0.0}, {0.5 * wb, 0.0}, {0.5 * wt, h}, {-0.5 * wt, h}};
  // d/dwt, d/dalpha (dcot/dalpha = -1/sin^2), d/dh
  const double dC[4][3][2] = {{{-0.5, 0}, {h * cs2, 0}, {-ct, 0}},
                              {{0.5, 0}, {-h * cs2, 0}, {ct, 0}},
                              {{0.5, 0}, {0, 0}, {0, 1}},
                              {{-0.5, 0}, {0, 0}, {0, 1}}};
  k %= gm.np;
  int a, n, i;
  if (k < gm.nb) {
    a = 0; n = gm.nb; i = k;
  } else if (k < gm.nb + gm.ns) {
    a = 1; n = gm.ns; i = k - gm.nb;
  } else if (k < gm.nb + gm.ns + gm.nt) {
    a = 2; n = gm.nt; i = k - gm.nb - gm.ns;
  } else {
    a = 3; n = gm.ns; i = k - gm.nb - gm.ns - gm.nt;
  }
  const int b = (a + 1) & 3;
  const double u = (double)i / n;
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    P[e] = (1.0 - u) * C[a][e] + u * C[b][e];
#pragma unroll
    for (int j = 0; j < 3; ++j) dP[j][e] = (1.0 - u) * dC[a][j][e] + u * dC[b][j][e];
  }
}

// Triangle of slot s: vertices v[3][3] and dv[3 vertices][3 params][3]; returns false for padding.
__device__ inline bool g_triangle(const GMesh& gm, int64_t s, const double th[3], double v[3][3],
                                  double dv[3][3][3]) {
  const int B = (int)(s / 128), j = (int)(s % 128);
  double P0[2], P1[2], dP0[3][2], dP1[3][2];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int e = 0; e < 3; ++e) dv[a][p][e] = 0.0;
  if (B < gm.nlb) {
    const int per = gm.np / 8;
    const int bi = B / per, bk = B % per, qd = j / 2, li = qd / 8, lk = qd % 8;
    const int i = bi * 8 + li, k = bk * 8 + lk;
    const double x0 = -0.5 * gm.L + gm.L * i / gm.nx, x1 = -0.5 * gm.L + gm.L * (i + 1) / gm.nx;
    g_perim(gm, k, th, P0, dP0);
    g_perim(gm, k + 1, th, P1, dP1);
    // quad a = (x0, P_k), b = (x0, P_k+1), c = (x1, P_k+1), d = (x1, P_k); triangles (a,b,c), (a,c,d)
    const double xs[4] = {x0, x0, x1, x1};
    const int ps[4] = {0, 1, 1, 0};
    const int idx[2][3] = {{0, 1, 2}, {0, 2, 3}};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int q = idx[j & 1][a];
      const double* P = ps[q] ? P1 : P0;
      v[a][0] = xs[q];
      v[a][1] = P[0];
      v[a][2] = P[1];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        dv[a][p][1] = ps[q] ? dP1[p][0] : dP0[p][0];
        dv[a][p][2] = ps[q] ? dP1[p][1] : dP0[p][1];
      }
    }
    return true;
  }
  const int per = gm.np / 16;
  const int B2 = B - gm.nlb, ci = B2 / per, gi = B2 % per;
  const double x = ci ? 0.5 * gm.L : -0.5 * gm.L;
  // cap centre: mean of the corners = (0, h / 2)
  const double M[3] = {x, 0.0, 0.5 * th[2]};
  if (j >= 16) {  // padding: zero area at the cap centre
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int e = 0; e < 3; ++e) v[a][e] = M[e];
    return false;
  }
  const int k = gi * 16 + j;
  g_perim(gm, k, th, P0, dP0);
  g_perim(gm, k + 1, th, P1, dP1);
  // +x cap: (M, P_k, P_k+1); -x cap: (M, P_k+1, P_k): outward normals
  const double* Pa = ci ? P0 : P1;
  const double* Pb = ci ? P1 : P0;
  v[0][0] = M[0];
  v[0][1] = M[1];
  v[0][2] = M[2];
  dv[0][2][2] = 0.5;  // dM_z / dh
  v[1][0] = x;
  v[1][1] = Pa[0];
  v[1][2] = Pa[1];
  v[2][0] = x;
  v[2][1] = Pb[0];
  v[2][2] = Pb[1];
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    dv[1][p][1] = ci ? dP0[p][0] : dP1[p][0];
    dv[1][p][2] = ci ? dP0[p][1] : dP1[p][1];
    dv[2][p][1] = ci ? dP1[p][0] : dP0[p][0];
    dv[2][p][2] = ci ? dP1[p][1] : dP0[p][1];
  }
  return true;
}

__device__ inline void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// rec64 [slot][8] = (A, c, 0, 0) with A = (v2 - v1) x (v3 - v1) / 2; drec [slot][3][6] = (dA, dc)
__global__ __launch_bounds__(256) void k_grating_mesh(GMesh gm, const GState* __restrict__ st, int cand0,
                                                      double* __restrict__ rec64, double* __restrict__ drec) {
  const int item = blockIdx.y;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= gm.nslots) return;
  const double* th = st[cand0 + item].trial;
  double v[3][3], dv[3][3][3];
  const bool real = g_triangle(gm, s, th, v, dv);
  double e1[3], e2[3], A[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    e1[k] = v[1][k] - v[0][k];
    e2[k] = v[2][k] - v[0][k];
  }
  cross3(e1, e2, A);
  double* r = rec64 + (int64_t)item * gm.nslots * 8 + 8 * s;
  double* dr = drec + ((int64_t)item * gm.nslots + s) * 18;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    r[k] = real ? 0.5 * A[k] : 0.0;
    r[3 + k] = (v[0][k] + v[1][k] + v[2][k]) / 3.0;
  }
  r[6] = r[7] = 0.0;
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    double de1[3], de2[3], t1[3], t2[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      de1[k] = dv[1][p][k] - dv[0][p][k];
      de2[k] = dv[2][p][k] - dv[0][p][k];
    }
    cross3(de1, e2, t1);
    cross3(e1, de2, t2);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dr[6 * p + k] = real ? 0.5 * (t1[k] + t2[k]) : 0.0;
      dr[6 * p + 3 + k] = real ? (dv[0][p][k] + dv[1][p][k] + dv[2][p][k]) / 3.0 : 0.0;
    }
  }
}

// Block origins (mean centroid of the block's real triangles, fixed-order LDS tree) and the FP32
// block-relative records the forward kernel reads.
// Block origins and FP32 records of every block; also each block's largest |A_t| and |c_t| (the
// split-fp16 forward's operand scales need bounds on both), reduced per item by k_grating_scales.
__global__ __launch_bounds__(128) void k_grating_blocks(GMesh gm, const double* __restrict__ rec64,
                                                        float* __restrict__ rec32, double* __restrict__ borig,
                                                        float* __restrict__ bmax) {
  __shared__ double sh[6][128];
  const int item = blockIdx.y, B = blockIdx.x, t = threadIdx.x;
  const int64_t s = (int64_t)B * 128 + t;
  const double* r = rec64 + ((int64_t)item * gm.nslots + s) * 8;
  const bool real = r[0] != 0.0 || r[1] != 0.0 || r[2] != 0.0;
  sh[0][t] = real ? r[3] : 0.0;
  sh[1][t] = real ? r[4] : 0.0;
  sh[2][t] = real ? r[5] : 0.0;
  sh[3][t] = real ? 1.0 : 0.0;
  sh[4][t] = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  sh[5][t] = real ? sqrt(r[3] * r[3] + r[4] * r[4] + r[5] * r[5]) : 0.0;
  __syncthreads();
  for (int w = 64; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int k = 0; k < 4; ++k) sh[k][t] += sh[k][t + w];
      sh[4][t] = fmax(sh[4][t], sh[4][t + w]);
      sh[5][t] = fmax(sh[5][t], sh[5][t + w]);
    }
    __syncthreads();
  }
  const double n = sh[3][0] > 0 ? sh[3][0] : 1.0;
  const double o[3] = {sh[0][0] / n, sh[1][0] / n, sh[2][0] / n};
  if (t == 0) {
    double* bo = borig + ((int64_t)item * gm.nblocks + B) * 4;
    bo[0] = o[0];
    bo[1] = o[1];
    bo[2] = o[2];
    bo[3] = 0.0;
    // raised by 1e-6 relative, so the bounds still hold after rounding to FP32
    float* bm = bmax + ((int64_t)item * gm.nblocks + B) * 2;
    bm[0] = (float)(sh[4][0] * (1.0 + 1e-6));
    bm[1] = (float)(sh[5][0] * (1.0 + 1e-6));
  }
  float* f = rec32 + ((int64_t)item * gm.nslots + s) * 8;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    f[k] = (float)r[k];
    f[3 + k] = real ? (float)(r[3 + k] - o[k]) : 0.0f;
  }
  f[6] = f[7] = 0.0f;
}

// Largest |s n| and |c| over each item's blocks into its product-kernel item (one workgroup per item).
__global__ __launch_bounds__(256) void k_grating_scales(int nblocks, const float* __restrict__ bmax,
                                                        ProductItem* __restrict__ pitems) {
  __shared__ float sh[2][256];
  const int item = blockIdx.x, t = threadIdx.x;
  float a = 0.0f, c = 0.0f;
  for (int b = t; b < nblocks; b += 256) {
    a = fmaxf(a, bmax[((int64_t)item * nblocks + b) * 2]);
    c = fmaxf(c, bmax[((int64_t)item * nblocks + b) * 2 + 1]);
  }
  sh[0][t] = a;
  sh[1][t] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      sh[0][t] = fmaxf(sh[0][t], sh[0][t + w]);
      sh[1][t] = fmaxf(sh[1][t], sh[1][t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    pitems[item].snmax = sh[0][0];
    pitems[item].rmax = sh[1][0];
  }
}

// ----- pixel pass: amplitude, intensity, chi^2, adjoint weights --------------------------------

__device__ inline double g_dirichlet(double x, int N) {
  // sin(N x / 2) / sin(x / 2) on the principal branch of x; limit N at x -> 0
  x -= kTwoPi * rint(x * kInvTwoPi);
  const double sh = sin(0.5 * x);
  if (fabs(sh) < 1e-12) return (double)N;
  return sin(0.5 * N * x) / sh;
}

// One thread per pixel, blockIdx.y = item. Reads the raw sums (chunked or not) and forms the pixel's
// model intensity. STATS: the six masked sums of the normalised metrics as per-block partials (spart).
// Otherwise: Wt = -i W / |q|^2 (FP32) for the adjoint, with W the derivative of the objective through
// the pixel's amplitude (seed dD/dI from metric_seed, constants per candidate in mconst), per-block
// partial sums of the objective and dD/ds, and optionally I.
template <bool STATS>
__global__ __launch_bounds__(256) void k_grating_pixel(GPixArgs a, const GState* __restrict__ st, int cand0,
                                                       const double* __restrict__ fbuf, int nchunks, int64_t nq,
                                                       int metric, const double* __restrict__ mconst,
                                                       float* __restrict__ wt, double* __restrict__ cpart,
                                                       double* __restrict__ Iout) {
  constexpr int NS = STATS ? 6 : 2;
  __shared__ double sh[NS][256];
  const int item = blockIdx.y;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const GState& S = st[cand0 + item];
  const double wt_t = S.trial[0], al = S.trial[1], h = S.trial[2], sc = S.trial[3];
  // radius of the line about its centre, for the small-|q| exclusion
  const double wb = wt_t + 2.0 * h * cos(al) / sin(al);
  const double Rl = sqrt(0.25 * a.L * a.L + 0.25 * fmax(wb, wt_t) * fmax(wb, wt_t) + 0.25 * h * h);
  double acc[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) acc[k] = 0.0;
  if (p < a.npix) {
    const double* fb = fbuf + (int64_t)item * nchunks * nq * 2;
    cd Fb[4] = {};
    double q2b[4] = {0, 0, 0, 0};
    bool near0 = false;
    cd A = cd{0.0, 0.0};
    const double qx = a.qx[p], qy = a.qy[p];
    for (int b = 0; b < a.nbranch; ++b) {
      const int64_t q = b * a.npix + p;
      cd Sb = cd{0.0, 0.0};
      for (int c = 0; c < nchunks; ++c) Sb = Sb + cd{fb[2 * (c * nq + q)], fb[2 * (c * nq + q) + 1]};
      const double zr = a.qz[2 * q], zi = a.qz[2 * q + 1];
      const double q2 = qx * qx + qy * qy + zr * zr + zi * zi;
      q2b[b] = q2;
      near0 = near0 || sqrt(q2) * Rl < kSeriesQR;
      Fb[b] = q2 > 0.0 ? mul_mi(Sb) * (1.0 / q2) : cd{0.0, 0.0};
      const cd fcb = a.fc ? cd{a.fc[2 * q], a.fc[2 * q + 1]} : cd{1.0, 0.0};
      A = A + fcb * Fb[b];
    }
    const cd dn2 = cd{a.dn2r, a.dn2i};
    A = dn2 * A;
    const double sf = g_dirichlet(qy * a.pitch, a.periods);
    const double I = sc * sf * sf * norm2(A);
    const double O = a.obs[p];
    const double m = (a.mask ? a.mask[p] : 1.0) * (near0 ? 0.0 : 1.0);
    if (STATS) {
      const double sI = I > 0.0 ? sqrt(I) : 0.0, sO = O > 0.0 ? sqrt(O) : 0.0;
      acc[0] = m * sI * sI;
      acc[1] = m * sI * sO;
      acc[2] = m * sO * sO;
      acc[3] = m * m * sI * sI;
      acc[4] = m * m * sI * sO;
      acc[5] = m * m * sO * sO;
    } else {
      if (Iout) Iout[(int64_t)item * a.npix + p] = I;
      double k[3] = {0.0, 0.0, 0.0};
      if (mconst) {
        k[0] = mconst[4 * item];
        k[1] = mconst[4 * item + 1];
        k[2] = mconst[4 * item + 2];
      }
      double term;
      const double gI = metric_seed(metric, I, O, m, k, &term);
      acc[0] = term;
      acc[1] = sc != 0.0 ? gI * I / sc : 0.0;
      // W_b = 2 gI s SF^2 conj(A) dn2 fc_b; Wt_b = -i W_b / |q_b|^2
      const cd base = (2.0 * gI * sc * sf * sf) * (conj(A) * dn2);
      for (int b = 0; b < a.nbranch; ++b) {
        const int64_t q = b * a.npix + p;
        const cd fcb = a.fc ? cd{a.fc[2 * q], a.fc[2 * q + 1]} : cd{1.0, 0.0};
        const cd W = base * fcb;
        const double s2 = q2b[b] > 0.0 ? 1.0 / q2b[b] : 0.0;
        float* o = wt + ((int64_t)item * nq + q) * 2;
        o[0] = (float)(W.i * s2);
        o[1] = (float)(-W.r * s2);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < NS; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < NS) {
    double* o = cpart + ((int64_t)item * gridDim.x + blockIdx.x) * NS;
    o[threadIdx.x] = sh[threadIdx.x][0];
  }
}

// Per candidate: the six sums of the statistics pass (fixed block order) -> the normalised metric's
// distance and seed constants, mconst[item] = (k0, k1, k2, D).
__global__ void k_grating_mconst(int gsize, int metric, const double* __restrict__ spart, int nblk,
                                 double* __restrict__ mconst) {
  const int item = blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= gsize) return;
  double S[6] = {0, 0, 0, 0, 0, 0};
  for (int b = 0; b < nblk; ++b)
#pragma unroll
    for (int k = 0; k < 6; ++k) S[k] += spart[((int64_t)item * nblk + b) * 6 + k];
  double kk[3];
  const double D = metric_from_sums(metric, S, kk);
  mconst[4 * item] = kk[0];
  mconst[4 * item + 1] = kk[1];
  mconst[4 * item + 2] = kk[2];
  mconst[4 * item + 3] = D;
}

// ----- gradient contraction --------------------------------------------------------------------

// Per slot: (dL/dA, dL/dc) from the adjoint forms, dotted with (dA, dc)/dtheta; per-block sums.
__global__ __launch_bounds__(256) void k_grating_grad(GMesh gm, const double* __restrict__ apart, int nyt,
                                                      const double* __restrict__ rec64,
                                                      const double* __restrict__ drec,
                                                      const double* __restrict__ qxp, double* __restrict__ gpart) {
  __shared__ double sh[3][256];
  const int item = blockIdx.y;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double gj[3] = {0, 0, 0};
  if (s < gm.nslots) {
    const double* rec = rec64 + (int64_t)item * gm.nslots * 8;
    const double* part = apart + (int64_t)item * 6 * nyt * gm.nslots * 2;
    double gA[3], gc[3];
    adj_assemble(part, nyt, gm.nslots, s, rec, qxp[0], gA, gc);
    const double* dr = drec + ((int64_t)item * gm.nslots + s) * 18;
#pragma unroll
    for (int p = 0; p < 3; ++p)
      gj[p] = gA[0] * dr[6 * p] + gA[1] * dr[6 * p + 1] + gA[2] * dr[6 * p + 2] + gc[0] * dr[6 * p + 3] +
              gc[1] * dr[6 * p + 4] + gc[2] * dr[6 * p + 5];
  }
#pragma unroll
  for (int p = 0; p < 3; ++p) sh[p][threadIdx.x] = gj[p];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int p = 0; p < 3; ++p) sh[p][threadIdx.x] += sh[p][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* o = gpart + ((int64_t)item * gridDim.x + blockIdx.x) * 3;
#pragma unroll
    for (int p = 0; p < 3; ++p) o[p] = sh[p][0];
  }
}

// chi^2 and gradient of each item at its trial point (fixed-order sums of the block partials).
__global__ void k_grating_reduce(GState* __restrict__ st, int cand0, int gsize, const double* __restrict__ cpart,
                                 int ncb, const double* __restrict__ gpart, int ngb,
                                 const double* __restrict__ mconst) {
  const int item = blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= gsize) return;
  double c2 = 0.0, cs = 0.0, g[3] = {0, 0, 0};
  for (int b = 0; b < ncb; ++b) {
    c2 += cpart[((int64_t)item * ncb + b) * 2];
    cs += cpart[((int64_t)item * ncb + b) * 2 + 1];
  }
  for (int b = 0; b < ngb; ++b)
#pragma unroll
    for (int p = 0; p < 3; ++p) g[p] += gpart[((int64_t)item * ngb + b) * 3 + p];
  GState& S = st[cand0 + item];
  S.ft = mconst ? mconst[4 * item + 3] : c2;
  S.gt[0] = g[0];
  S.gt[1] = g[1];
  S.gt[2] = g[2];
  S.gt[3] = cs;
}

// ----- optimiser ---------------------------------------------------------------------------------

__global__ void k_grating_begin_eval(GState* st, int n) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  for (int k = 0; k < kGPar; ++k) st[c].trial[k] = st[c].th[k];
}

__global__ void k_grating_end_eval(GState* st, int n) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  st[c].f = st[c].ft;
  for (int k = 0; k < kGPar; ++k) st[c].g[k] = st[c].gt[k];
}

// Keeps the trapezoid valid: positive top width and height, 0.1 <= alpha <= pi/2 (no overhang, so
// the bottom width wb = wt + 2 h cot(alpha) >= wt), positive scale.
__device__ inline void g_clamp(double th[kGPar]) {
  th[0] = fmax(th[0], 0.1);
  th[1] = fmin(fmax(th[1], 0.1), 1.5707963267948966);
  th[2] = fmax(th[2], 0.1);
  th[3] = fmax(th[3], 1e-30);
}

// One L-BFGS step per candidate in scaled coordinates u = theta / ref: accept the trial point when it
// satisfies the Armijo condition (then update the memory and take the two-loop direction with step
// 1), otherwise halve the step along the same direction. Work per iteration is fixed: one chi^2 +
// gradient evaluation, so the whole iteration replays from one graph.
__global__ void k_grating_lbfgs(GState* st, int n) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  GState& S = st[c];
  if (S.status) return;  // converged or stalled: keep the accepted point
  double gu[kGPar];
  const bool finite = isfinite(S.ft);
  if (S.iter == 0) {
    S.f = S.ft;
    for (int k = 0; k < kGPar; ++k) {
      S.th[k] = S.trial[k];
      S.g[k] = S.gt[k];
    }
    double nrm = 0.0;
    for (int k = 0; k < kGPar; ++k) {
      gu[k] = S.g[k] * S.ref[k];
      nrm += gu[k] * gu[k];
    }
    nrm = sqrt(nrm);
    for (int k = 0; k < kGPar; ++k) S.d[k] = nrm > 0 ? -gu[k] / nrm : 0.0;
    S.step = 0.02;  // 2 % of the parameter scale
  } else {
    double slope = 0.0;
    for (int k = 0; k < kGPar; ++k) slope += S.g[k] * S.ref[k] * S.d[k];
    if (finite && S.ft <= S.f + 1e-4 * S.step * slope) {
      double s[kGPar], y[kGPar], sy = 0.0, yy = 0.0, ss = 0.0;
      for (int k = 0; k < kGPar; ++k) {
        s[k] = (S.trial[k] - S.th[k]) / S.ref[k];
        y[k] = (S.gt[k] - S.g[k]) * S.ref[k];
        sy += s[k] * y[k];
        yy += y[k] * y[k];
        ss += s[k] * s[k];
      }
      if (sy > 1e-12 * sqrt(ss * yy)) {
        const int slot = S.head;
        for (int k = 0; k < kGPar; ++k) {
          S.S[slot][k] = s[k];
          S.Y[slot][k] = y[k];
        }
        S.rho[slot] = 1.0 / sy;
        S.head = (S.head + 1) % kLbfgsMem;
        S.nhist = min(S.nhist + 1, kLbfgsMem);
      }
      S.f = S.ft;
      for (int k = 0; k < kGPar; ++k) {
        S.th[k] = S.trial[k];
        S.g[k] = S.gt[k];
      }
      // two-loop recursion on the scaled gradient
      double q[kGPar], alpha[kLbfgsMem];
      for (int k = 0; k < kGPar; ++k) q[k] = S.g[k] * S.ref[k];
      for (int i = 0; i < S.nhist; ++i) {
        const int j = (S.head - 1 - i + kLbfgsMem) % kLbfgsMem;
        double a = 0.0;
        for (int k = 0; k < kGPar; ++k) a += S.S[j][k] * q[k];
        a *= S.rho[j];
        alpha[i] = a;
        for (int k = 0; k < kGPar; ++k) q[k] -= a * S.Y[j][k];
      }
      double gamma = 1e-2;
      if (S.nhist > 0) {
        const int j = (S.head - 1 + kLbfgsMem) % kLbfgsMem;
        double yy2 = 0.0;
        for (int k = 0; k < kGPar; ++k) yy2 += S.Y[j][k] * S.Y[j][k];
        gamma = yy2 > 0 ? 1.0 / (S.rho[j] * yy2) : gamma;
      }
      for (int k = 0; k < kGPar; ++k) q[k] *= gamma;
      for (int i = S.nhist - 1; i >= 0; --i) {
        const int j = (S.head - 1 - i + kLbfgsMem) % kLbfgsMem;
        double b = 0.0;
        for (int k = 0; k < kGPar; ++k) b += S.Y[j][k] * q[k];
        b *= S.rho[j];
        for (int k = 0; k < kGPar; ++k) q[k] += S.S[j][k] * (alpha[i] - b);
      }
      double dg = 0.0;
      for (int k = 0; k < kGPar; ++k) {
        S.d[k] = -q[k];
        dg += S.d[k] * S.g[k] * S.ref[k];
      }
      if (!(dg < 0.0)) {  // not a descent direction: restart from the scaled gradient
        double nrm = 0.0;
        for (int k = 0; k < kGPar; ++k) nrm += (S.g[k] * S.ref[k]) * (S.g[k] * S.ref[k]);
        nrm = sqrt(nrm);
        for (int k = 0; k < kGPar; ++k) S.d[k] = nrm > 0 ? -S.g[k] * S.ref[k] / nrm * 0.02 : 0.0;
        S.nhist = 0;
      }
      S.step = 1.0;
    } else {
      S.step *= 0.5;
      if (S.step < 1e-10) S.status = 1;
    }
  }
  S.iter += 1;
  for (int k = 0; k < kGPar; ++k) S.trial[k] = S.th[k] + S.step * S.d[k] * S.ref[k];
  g_clamp(S.trial);
}

}  // namespace hgx

using namespace hgx;

// ----- plan --------------------------------------------------------------------------------------

namespace hgx {

struct GratingPlan {
  bool live = false;
  hgx_grating_model model{};
  GMesh gm{};
  GPixArgs pa{};
  int ncand = 0, G = 0, ngroups = 0;
  int nchunks = 1, nyt = 0, ntt = 0, upi_fwd = 0, upi_adj = 0;
  int npb = 0, ngb = 0;  // pixel blocks, gradient blocks per item
  int pmode = 0;         // product-kernel variant of the forward pass, fixed at creation
  int adj_variant = 2;   // adjoint kernel variant (0 FP32, 2/4 split fp16), fixed at creation
  bool adj_f16 = true;
  bool cq = false;
  GState* d_state = nullptr;
  double* d_rec64 = nullptr;
  double* d_drec = nullptr;
  float* d_rec32 = nullptr;
  double* d_borig = nullptr;
  float* d_bmax = nullptr;
  double* d_fbuf = nullptr;
  float* d_wt = nullptr;
  float* d_pmax = nullptr;
  uint4* d_w16 = nullptr;
  int* d_wexp = nullptr;
  double* d_apart = nullptr;
  double* d_gpart = nullptr;
  double* d_cpart = nullptr;
  double* d_spart = nullptr;   // statistics-pass partials [G][npb][6]
  double* d_mconst = nullptr;  // per-candidate metric constants [G][4]
  int metric = HGX_METRIC_CHI2;
  uint64_t grid_gen = 0;  // the context's grid generation the device tables were built for
  ProductItem* d_pitems = nullptr;
  ProductUnit* d_punits = nullptr;
  AdjItem* d_aitems = nullptr;
  hipGraphExec_t eval_exec = nullptr, step_exec = nullptr;
  hipGraph_t eval_graph = nullptr, step_graph = nullptr;
};

}  // namespace hgx

namespace {

std::vector<GratingPlan*>& gplans(Context* c) {
  static_assert(sizeof(GState) % 8 == 0, "state layout");
  return c->gratings;
}

void free_plan(GratingPlan& p) {
  if (p.eval_exec) (void)hipGraphExecDestroy(p.eval_exec);
  if (p.step_exec) (void)hipGraphExecDestroy(p.step_exec);
  if (p.eval_graph) (void)hipGraphDestroy(p.eval_graph);
  if (p.step_graph) (void)hipGraphDestroy(p.step_graph);
  void* ptrs[] = {p.d_state, p.d_rec64, p.d_drec, p.d_rec32, p.d_borig, p.d_bmax, p.d_fbuf, p.d_wt, p.d_pmax,
                  p.d_w16, p.d_wexp, p.d_apart, p.d_gpart, p.d_cpart, p.d_spart, p.d_mconst, p.d_pitems,
                  p.d_punits, p.d_aitems};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  p = GratingPlan();
}

// One group's evaluation (gsize candidates from cand0) on stream s. Returns nothing; launch errors
// are checked by the caller.
void group_eval(const GratingPlan& p, Context* c, hipStream_t s, int cand0, int gsize, bool want_I, double* Iout) {
  const GMesh& gm = p.gm;
  dim3 gmesh(ceil_div(gm.nslots, 256), gsize);
  k_grating_mesh<<<gmesh, 256, 0, s>>>(gm, p.d_state, cand0, p.d_rec64, p.d_drec);
  dim3 gblk(gm.nblocks, gsize);
  k_grating_blocks<<<gblk, 128, 0, s>>>(gm, p.d_rec64, p.d_rec32, p.d_borig, p.d_bmax);
  // the meshes are built here, so the split-fp16 operand scales' bounds are too
  k_grating_scales<<<gsize, 256, 0, s>>>(gm.nblocks, p.d_bmax, p.d_pitems);
  product_launch_units(s, p.d_pitems, p.d_punits, gsize * p.upi_fwd, p.cq, p.pmode);
  dim3 gpix(p.npb, gsize);
  const bool norm = p.metric == HGX_METRIC_SQRT_UNIT_L2 || p.metric == HGX_METRIC_SQRT_C_L2;
  if (norm && !want_I) {
    k_grating_pixel<true><<<gpix, 256, 0, s>>>(p.pa, p.d_state, cand0, p.d_fbuf, p.nchunks, c->nq, p.metric,
                                               nullptr, nullptr, p.d_spart, nullptr);
    k_grating_mconst<<<ceil_div(gsize, 64), 64, 0, s>>>(gsize, p.metric, p.d_spart, p.npb, p.d_mconst);
  }
  k_grating_pixel<false><<<gpix, 256, 0, s>>>(p.pa, p.d_state, cand0, p.d_fbuf, p.nchunks, c->nq, p.metric,
                                              norm && !want_I ? p.d_mconst : nullptr, p.d_wt, p.d_cpart,
                                              want_I ? Iout : nullptr);
  if (want_I) return;
  AdjItem none{};
  if (p.adj_f16) {
    const int nz = c->nbranch * c->nrow;
    adjoint_split_weights(s, gsize, p.d_wt, 2 * c->nq, nz, c->ncol, p.d_pmax, adjoint_pmax_count(nz, c->ncol), p.d_w16,
                          kAdjW16Planes * (int64_t)ceil_div(nz, 8) * c->ncol, p.d_wexp, c->ncol);
  }
  adjoint_product_launch(s, p.d_aitems, none, p.upi_adj, gsize * p.upi_adj, p.cq, p.adj_variant);
  dim3 ggrad(p.ngb, gsize);
  k_grating_grad<<<ggrad, 256, 0, s>>>(gm, p.d_apart, p.nyt, p.d_rec64, p.d_drec, c->d_qx, p.d_gpart);
  k_grating_reduce<<<ceil_div(gsize, 64), 64, 0, s>>>(p.d_state, cand0, gsize, p.d_cpart, p.npb, p.d_gpart, p.ngb,
                                                       norm ? p.d_mconst : nullptr);
}

void all_groups(const GratingPlan& p, Context* c, hipStream_t s) {
  for (int g = 0; g < p.ngroups; ++g) {
    const int cand0 = g * p.G;
    group_eval(p, c, s, cand0, std::min(p.G, p.ncand - cand0), false, nullptr);
  }
}

int capture(Context* c, GratingPlan& p, bool step) {
  hipStream_t cap;
  HGX_CHECK_HIP(c, hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  HGX_CHECK_HIP(c, hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
  if (!step) k_grating_begin_eval<<<ceil_div(p.ncand, 64), 64, 0, cap>>>(p.d_state, p.ncand);
  all_groups(p, c, cap);
  if (step) k_grating_lbfgs<<<ceil_div(p.ncand, 64), 64, 0, cap>>>(p.d_state, p.ncand);
  else k_grating_end_eval<<<ceil_div(p.ncand, 64), 64, 0, cap>>>(p.d_state, p.ncand);
  hipGraph_t gr;
  hipError_t e = hipStreamEndCapture(cap, &gr);
  (void)hipStreamDestroy(cap);
  if (e != hipSuccess) return set_error(c, HGX_ERR_DEVICE, "grating: capture failed: %s", hipGetErrorString(e));
  hipGraphExec_t ex;
  HGX_CHECK_HIP(c, hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
  (step ? p.step_graph : p.eval_graph) = gr;
  (step ? p.step_exec : p.eval_exec) = ex;
  return HGX_OK;
}

GratingPlan* get_plan(Context* c, int id) {
  auto& P = gplans(c);
  if (id < 0 || id >= (int)P.size() || !P[id] || !P[id]->live) return nullptr;
  return P[id];
}

// A plan's tables and captured graphs hold the grid's device addresses and sizes of creation time.
int check_grid(Context* c, const GratingPlan& p, const char* who) {
  if (c->grid_gen != p.grid_gen)
    return set_error(c, HGX_ERR_INVALID_ARG, "%s: the q-grid was replaced or extended after the plan was "
                     "created; create the plan again", who);
  return HGX_OK;
}

}  // namespace

namespace hgx {
void free_grating_plans(Context* c) {
  for (auto* p : c->gratings)
    if (p) {
      free_plan(*p);
      delete p;
    }
  c->gratings.clear();
}
}  // namespace hgx

extern "C" int hgx_grating_plan_create(hgx_ctx* h, const hgx_grating_model* model, int ncand, const double* fc,
                                       const double* I_obs, const double* mask, int* plan) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!model || !I_obs || !plan || ncand < 1) return set_error(c, HGX_ERR_INVALID_ARG, "grating_plan_create: bad arguments");
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "grating_plan_create: no q-grid");
  if (!c->product) return set_error(c, HGX_ERR_UNSUPPORTED, "grating_plan_create: needs a product q-grid");
  const hgx_grating_model& m = *model;
  const int np = m.n_bottom + 2 * m.n_side + m.n_top;
  if (m.n_bottom < 1 || m.n_side < 1 || m.n_top < 1 || np % 16 != 0 || m.n_length < 8 || m.n_length % 8 != 0 ||
      !(m.length > 0) || !(m.pitch > 0) || m.periods < 1)
    return set_error(c, HGX_ERR_INVALID_ARG,
                     "grating_plan_create: need perimeter segments %% 16 == 0, n_length %% 8 == 0, positive sizes");
  if (int rc = set_device(c)) return rc;
  GratingPlan p;
  p.model = m;
  p.ncand = ncand;
  GMesh& gm = p.gm;
  gm.L = m.length;
  gm.nb = m.n_bottom;
  gm.ns = m.n_side;
  gm.nt = m.n_top;
  gm.nx = m.n_length;
  gm.np = np;
  gm.nlb = (m.n_length / 8) * (np / 8);
  gm.ncb = 2 * (np / 16);
  gm.nblocks = gm.nlb + gm.ncb;
  gm.nslots = (int64_t)gm.nblocks * kBlockTri;
  GPixArgs& pa = p.pa;
  pa.qx = c->d_qx;
  pa.qy = c->d_qy;
  pa.qz = c->d_qz;
  pa.npix = c->npix;
  pa.nbranch = c->nbranch;
  pa.fc = fc;
  pa.obs = I_obs;
  pa.mask = mask;
  pa.dn2r = m.dn2_re;
  pa.dn2i = m.dn2_im;
  pa.pitch = m.pitch;
  pa.periods = m.periods;
  pa.L = m.length;
  p.cq = !c->qz_real;
  const int nz = c->nbranch * c->nrow;
  p.pmode = product_mode(c);
  const int nzt = (int)ceil_div(nz, product_zrows(p.pmode)), nyt_f = (int)ceil_div(c->ncol, product_ycols(p.pmode));
  const int64_t bchunk = std::min<int64_t>(kGChunkBlocks, gm.nblocks);
  p.nchunks = (int)((gm.nblocks + bchunk - 1) / bchunk);
  p.nyt = (int)ceil_div(c->ncol, 32);
  p.ntt = (int)(gm.nslots / 32);
  p.upi_fwd = nzt * nyt_f * p.nchunks;
  p.adj_variant = adjoint_product_variant(c);
  p.adj_f16 = p.adj_variant > 0;
  const int nzo = (int)ceil_div(nz, 8);
  {
    AdjItem shape{};
    shape.nblocks = gm.nblocks;
    shape.ncol = c->ncol;
    shape.nyt = p.nyt;
    const int64_t upi = adjoint_product_units(shape, p.adj_variant);
    if (upi > 2000000000) return set_error(c, HGX_ERR_INVALID_ARG, "grating_plan_create: problem too large");
    p.upi_adj = (int)upi;
  }
  p.npb = (int)ceil_div(c->npix, 256);
  p.ngb = (int)ceil_div(gm.nslots, 256);
  // group size: per-candidate device memory within a 24 GiB budget
  const double per = (double)gm.nslots * (8 * 8 + 18 * 8 + 8 * 4) + (double)gm.nblocks * 40 +
                     (double)p.nchunks * c->nq * 16 + (double)c->nq * 8 + 6.0 * p.nyt * gm.nslots * 16 +
                     (p.adj_f16 ? 96.0 * nzo * c->ncol + 4.0 * adjoint_pmax_count(nz, c->ncol) + 4.0 * c->ncol : 0.0) +
                     (double)(p.npb * 8 + p.ngb * 3 + 4) * 8;
  p.G = (int)std::max(1.0, std::min((double)ncand, 24.0 * (1 << 30) / per));
  p.ngroups = (ncand + p.G - 1) / p.G;
  if ((double)p.G * p.upi_fwd > 2.0e9 || (double)p.G * p.upi_adj > 2.0e9)
    return set_error(c, HGX_ERR_INVALID_ARG, "grating_plan_create: problem too large");
  const int G = p.G;
  auto fail = [&](const char* what) {
    (void)hipGetLastError();
    free_plan(p);
    return set_error(c, HGX_ERR_OOM, "grating_plan_create: allocation of %s failed", what);
  };
  if (hipMalloc(&p.d_state, sizeof(GState) * ncand) != hipSuccess) return fail("state");
  if (hipMalloc(&p.d_rec64, sizeof(double) * 8 * gm.nslots * G) != hipSuccess) return fail("records");
  if (hipMalloc(&p.d_drec, sizeof(double) * 18 * gm.nslots * G) != hipSuccess) return fail("record derivatives");
  if (hipMalloc(&p.d_rec32, sizeof(float) * 8 * gm.nslots * G) != hipSuccess) return fail("FP32 records");
  if (hipMalloc(&p.d_borig, sizeof(double) * 4 * gm.nblocks * G) != hipSuccess) return fail("block origins");
  if (hipMalloc(&p.d_bmax, sizeof(float) * 2 * gm.nblocks * G) != hipSuccess) return fail("block bounds");
  if (hipMalloc(&p.d_fbuf, sizeof(double) * 2 * p.nchunks * c->nq * G) != hipSuccess) return fail("sums");
  if (hipMalloc(&p.d_wt, sizeof(float) * 2 * c->nq * G) != hipSuccess) return fail("weights");
  if (p.adj_f16) {
    if (hipMalloc(&p.d_pmax, sizeof(float) * adjoint_pmax_count(nz, c->ncol) * G) != hipSuccess) return fail("weight maxima");
    if (hipMalloc(&p.d_w16, sizeof(uint4) * kAdjW16Planes * (size_t)nzo * c->ncol * G) != hipSuccess) return fail("split weights");
    if (hipMalloc(&p.d_wexp, sizeof(int) * (size_t)c->ncol * G) != hipSuccess) return fail("weight scales");
  }
  if (hipMalloc(&p.d_apart, sizeof(double) * 2 * 6 * p.nyt * gm.nslots * G) != hipSuccess) return fail("adjoint");
  if (hipMalloc(&p.d_gpart, sizeof(double) * 3 * p.ngb * G) != hipSuccess) return fail("gradient partials");
  if (hipMalloc(&p.d_cpart, sizeof(double) * 2 * p.npb * G) != hipSuccess) return fail("chi2 partials");
  if (hipMalloc(&p.d_spart, sizeof(double) * 6 * p.npb * G) != hipSuccess) return fail("metric partials");
  if (hipMalloc(&p.d_mconst, sizeof(double) * 4 * G) != hipSuccess) return fail("metric constants");
  std::vector<ProductItem> pit(G);
  std::vector<AdjItem> ait(G);
  for (int i = 0; i < G; ++i) {
    ProductItem& q = pit[i];
    q.qx = c->d_qx;
    q.qy = c->d_qy;
    q.qz = c->d_qz;
    q.rec = p.d_rec32 + (int64_t)i * gm.nslots * 8;
    q.borig = p.d_borig + (int64_t)i * gm.nblocks * 4;
    q.nblocks = gm.nblocks;
    q.bchunk = bchunk;
    q.nq = c->nq;
    q.nchunks = p.nchunks;
    q.ncol = c->ncol;
    q.nz = nz;
    double* fb = p.d_fbuf + (int64_t)i * p.nchunks * c->nq * 2;
    q.part = fb;  // chunk c at fb + 2 c nq (the pixel pass reads them all)
    q.out = fb;   // one chunk: raw sums in place
    AdjItem& a = ait[i];
    a.qx = c->d_qx;
    a.qy = c->d_qy;
    a.qz = c->d_qz;
    a.wt = p.d_wt + (int64_t)i * c->nq * 2;
    a.rec = p.d_rec64 + (int64_t)i * gm.nslots * 8;
    a.rec32 = q.rec;
    a.borig = q.borig;
    a.part = p.d_apart + (int64_t)i * 6 * p.nyt * gm.nslots * 2;
    a.ntri = gm.nslots;
    a.nblocks = gm.nblocks;
    a.ncol = c->ncol;
    a.nz = nz;
    a.nyt = p.nyt;
    a.nzo = nzo;
    a.w16 = p.adj_f16 ? p.d_w16 + (int64_t)i * kAdjW16Planes * nzo * c->ncol : nullptr;
    a.wexp = p.adj_f16 ? p.d_wexp + (int64_t)i * c->ncol : nullptr;
  }
  // forward units, item-major so a short last group runs a prefix
  std::vector<ProductUnit> units;
  units.reserve((size_t)G * p.upi_fwd);
  for (int i = 0; i < G; ++i)
    for (int ch = 0; ch < p.nchunks; ++ch)
      for (int zt = 0; zt < nzt; ++zt)
        for (int yt = 0; yt < nyt_f; ++yt) units.push_back(ProductUnit{i, zt, yt, ch});
  if (hipMalloc(&p.d_pitems, sizeof(ProductItem) * G) != hipSuccess) return fail("items");
  if (hipMalloc(&p.d_punits, sizeof(ProductUnit) * units.size()) != hipSuccess) return fail("units");
  if (hipMalloc(&p.d_aitems, sizeof(AdjItem) * G) != hipSuccess) return fail("adjoint items");
  std::vector<GState> st(ncand);
  if (hipMemcpy(p.d_pitems, pit.data(), sizeof(ProductItem) * G, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p.d_punits, units.data(), sizeof(ProductUnit) * units.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p.d_aitems, ait.data(), sizeof(AdjItem) * G, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p.d_state, st.data(), sizeof(GState) * ncand, hipMemcpyHostToDevice) != hipSuccess) {
    hipError_t e = hipGetLastError();
    free_plan(p);
    return set_error(c, HGX_ERR_DEVICE, "grating_plan_create: copy to the device failed: %s", hipGetErrorString(e));
  }
  p.grid_gen = c->grid_gen;
  p.live = true;
  auto& P = gplans(c);
  for (size_t i = 0; i < P.size(); ++i)
    if (!P[i]) {
      P[i] = new GratingPlan(p);
      *plan = (int)i;
      return HGX_OK;
    }
  P.push_back(new GratingPlan(p));
  *plan = (int)P.size() - 1;
  return HGX_OK;
}

extern "C" int hgx_grating_plan_release(hgx_ctx* h, int plan) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  GratingPlan* p = get_plan(c, plan);
  if (!p) return set_error(c, HGX_ERR_BAD_HANDLE, "grating_plan_release: unknown plan %d", plan);
  if (int rc = set_device(c)) return rc;
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  free_plan(*p);
  delete p;
  gplans(c)[plan] = nullptr;
  return HGX_OK;
}

extern "C" int hgx_grating_plan_metric(hgx_ctx* h, int plan, int metric) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  GratingPlan* p = get_plan(c, plan);
  if (!p) return set_error(c, HGX_ERR_BAD_HANDLE, "grating_plan_metric: unknown plan %d", plan);
  if (metric < HGX_METRIC_CHI2 || metric > HGX_METRIC_SQRT_C_L2)
    return set_error(c, HGX_ERR_INVALID_ARG, "grating_plan_metric: unknown metric %d", metric);
  if (metric == p->metric) return HGX_OK;
  if (int rc = set_device(c)) return rc;
  // the captured graphs hold the old metric's launches: drop them (captured again on next use)
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  if (p->eval_exec) (void)hipGraphExecDestroy(p->eval_exec);
  if (p->step_exec) (void)hipGraphExecDestroy(p->step_exec);
  if (p->eval_graph) (void)hipGraphDestroy(p->eval_graph);
  if (p->step_graph) (void)hipGraphDestroy(p->step_graph);
  p->eval_exec = p->step_exec = nullptr;
  p->eval_graph = p->step_graph = nullptr;
  p->metric = metric;
  return HGX_OK;
}

extern "C" int hgx_grating_set_theta(hgx_ctx* h, int plan, const double* theta) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  GratingPlan* p = get_plan(c, plan);
  if (!p || !theta) return set_error(c, HGX_ERR_INVALID_ARG, "grating_set_theta: bad plan or theta");
  if (int rc = set_device(c)) return rc;
  std::vector<GState> st(p->ncand);
  for (int i = 0; i < p->ncand; ++i) {
    for (int k = 0; k < kGPar; ++k) {
      const double v = theta[kGPar * i + k];
      if (!std::isfinite(v)) return set_error(c, HGX_ERR_INVALID_ARG, "grating_set_theta: non-finite theta");
      st[i].th[k] = st[i].trial[k] = v;
      st[i].ref[k] = std::max(std::fabs(v), k == 1 ? 0.1 : 1e-12);
    }
  }
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  HGX_CHECK_HIP(c, hipMemcpy(p->d_state, st.data(), sizeof(GState) * p->ncand, hipMemcpyHostToDevice));
  return HGX_OK;
}

extern "C" int hgx_grating_eval(hgx_ctx* h, int plan, int use_graph) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  GratingPlan* p = get_plan(c, plan);
  if (!p) return set_error(c, HGX_ERR_BAD_HANDLE, "grating_eval: unknown plan %d", plan);
  if (int rc = check_grid(c, *p, "grating_eval")) return rc;
  if (int rc = set_device(c)) return rc;
  if (!use_graph) {
    LaunchTimer lt(c, true);
    k_grating_begin_eval<<<ceil_div(p->ncand, 64), 64, 0, c->stream>>>(p->d_state, p->ncand);
    all_groups(*p, c, c->stream);
    k_grating_end_eval<<<ceil_div(p->ncand, 64), 64, 0, c->stream>>>(p->d_state, p->ncand);
    HGX_CHECK_LAUNCH(c);
    return HGX_OK;
  }
  if (!p->eval_exec)
    if (int rc = capture(c, *p, false)) return rc;
  LaunchTimer lt(c, true);
  HGX_CHECK_HIP(c, hipGraphLaunch(p->eval_exec, c->stream));
  return HGX_OK;
}

extern "C" int hgx_grating_lbfgs(hgx_ctx* h, int plan, int iterations) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  GratingPlan* p = get_plan(c, plan);
  if (!p || iterations < 0) return set_error(c, HGX_ERR_INVALID_ARG, "grating_lbfgs: bad plan or count");
  if (int rc = check_grid(c, *p, "grating_lbfgs")) return rc;
  if (int rc = set_device(c)) return rc;
  if (!p->step_exec)
    if (int rc = capture(c, *p, true)) return rc;
  for (int i = 0; i < iterations; ++i) {
    LaunchTimer lt(c, true);
    HGX_CHECK_HIP(c, hipGraphLaunch(p->step_exec, c->stream));
  }
  return HGX_OK;
}

extern "C" int hgx_grating_get(hgx_ctx* h, int plan, double* theta, double* chi2, double* grad) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  GratingPlan* p = get_plan(c, plan);
  if (!p) return set_error(c, HGX_ERR_BAD_HANDLE, "grating_get: unknown plan %d", plan);
  if (int rc = set_device(c)) return rc;
  std::vector<GState> st(p->ncand);
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  HGX_CHECK_HIP(c, hipMemcpy(st.data(), p->d_state, sizeof(GState) * p->ncand, hipMemcpyDeviceToHost));
  for (int i = 0; i < p->ncand; ++i) {
    for (int k = 0; k < kGPar; ++k) {
      if (theta) theta[kGPar * i + k] = st[i].th[k];
      if (grad) grad[kGPar * i + k] = st[i].g[k];
    }
    if (chi2) chi2[i] = st[i].f;
  }
  return HGX_OK;
}

extern "C" int hgx_grating_intensity(hgx_ctx* h, int plan, int cand, double* I) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  GratingPlan* p = get_plan(c, plan);
  if (!p || !I || cand < 0 || cand >= p->ncand)
    return set_error(c, HGX_ERR_INVALID_ARG, "grating_intensity: bad plan, candidate or output");
  if (int rc = check_grid(c, *p, "grating_intensity")) return rc;
  if (int rc = set_device(c)) return rc;
  k_grating_begin_eval<<<ceil_div(p->ncand, 64), 64, 0, c->stream>>>(p->d_state, p->ncand);
  group_eval(*p, c, c->stream, cand, 1, true, I);
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}
