// ff_numeric.hip -- triangulated-mesh form factors on gfx950.
//
// Approximate (centroid) method, src/ff/cpu/ff_tri_cpu.cpp:200-246:
//     F(q) = -i / |q|^2 * sum_t (q . s_t n_t) exp(i q . c_t)
// Exact (flat triangle) method, src/ff/cpu/ff_tri_cpu.cpp:46-195.
//
// Work decomposition (both precisions, so the FP64/FP32 comparison is like for like):
//   * a q-tile is kWG threads x QPT q-points per thread (q-points strided by kWG: coalesced);
//   * a triangle chunk is a run of 128-triangle blocks of the Morton-sorted mesh;
//   * grid = (q-tiles, chunks). With one chunk the kernel finishes the q-points itself, otherwise
//     it writes FP64 partial sums [chunk][q] and k_ff_finish reduces them in a fixed order
//     (deterministic; no atomics).
//   * triangle records are wave-uniform: the loop index depends only on blockIdx and the loop
//     counter, so the records arrive through scalar loads into SGPRs and feed the VALU directly.
//
// FP32C path: per block, coordinates are FP32 relative to the block's FP64 origin o_b, the phase
// is carried in turns (q/2pi . (c - o_b)) and fed straight to the hardware v_sin/v_cos (which take
// turns and reduce the argument themselves); the block's FP32 sum (plain, or Kahan) is rotated by
// exp(i q . o_b) and added to an FP64 accumulator. For |q| R < kSeriesQR (R = mesh radius about
// its centroid) the result is replaced by the FP64 moment series of the same sum, which has no
// cancellation at small q.
#include "numeric_core.hpp"

namespace hgx {

// ----- FP64 plain approximate kernel ---------------------------------------------------------

template <int QPT, bool CQ>
__global__ __launch_bounds__(kWG) void k_ff_approx64(QView g, Rot R, const double* __restrict__ rec, int64_t ntri,
                                                     int64_t tchunk, int nchunks, double* __restrict__ part,
                                                     double* __restrict__ out, Epi e) {
  const int64_t qbase = (int64_t)blockIdx.x * (kWG * QPT) + threadIdx.x;
  const int chunk = blockIdx.y;
  const int64_t t0 = (int64_t)chunk * tchunk;
  const int64_t t1 = min(ntri, t0 + tchunk);
  cd mq[QPT][3];
  double ar[QPT], ai[QPT];
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    int64_t q = qbase + (int64_t)k * kWG;
    load_q<CQ>(g, R, q < g.nq ? q : g.nq - 1, mq[k]);
  }
  approx64_tris<QPT, CQ>(mq, rec, t0, t1, ar, ai);
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    int64_t q = qbase + (int64_t)k * kWG;
    if (q >= g.nq) continue;
    // raw sums; k_ff_finish applies the epilogue (in place when there is a single chunk)
    double* dst = nchunks == 1 ? out + 2 * q : part + 2 * ((int64_t)chunk * g.nq + q);
    dst[0] = ar[k];
    dst[1] = ai[k];
  }
}

// ----- FP32 compensated approximate kernel ----------------------------------------------------

template <int QPT, bool CQ, bool KAHAN>
__global__ __launch_bounds__(kWG) void k_ff_approx32(QView g, Rot R, const float* __restrict__ rec,
                                                     const double* __restrict__ borig, int64_t nblocks,
                                                     int64_t bchunk, int nchunks, double* __restrict__ part,
                                                     double* __restrict__ out, Epi e) {
  const int64_t qbase = (int64_t)blockIdx.x * (kWG * QPT) + threadIdx.x;
  const int chunk = blockIdx.y;
  const int64_t b0 = (int64_t)chunk * bchunk;
  const int64_t b1 = min(nblocks, b0 + bchunk);
  double ar[QPT], ai[QPT];
  approx32_blocks<QPT, CQ, KAHAN>(g, R, qbase, rec, borig, b0, b1, ar, ai);
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    int64_t q = qbase + (int64_t)k * kWG;
    if (q >= g.nq) continue;
    // raw sums; k_ff_finish applies the epilogue (in place when there is a single chunk)
    double* dst = nchunks == 1 ? out + 2 * q : part + 2 * ((int64_t)chunk * g.nq + q);
    dst[0] = ar[k];
    dst[1] = ai[k];
  }
}

// Reduces partial sums over chunks in chunk order (or takes the raw sum the hot kernel left in
// `out` when there is one chunk) and applies the per-q epilogue, including the small-|q| series.
__global__ __launch_bounds__(kWG) void k_ff_finish(QView g, Rot R, int nchunks, const double* __restrict__ part,
                                                   const float* __restrict__ part32, double* __restrict__ out, Epi e) {
  int64_t q = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (q >= g.nq) return;
  cd mq[3];
  load_q(g, R, q, mq);
  const cd raw = nchunks == 1 ? cd{out[2 * q], out[2 * q + 1]} : sum_chunks(nchunks, part, part32, g.nq, q);
  cd F = finish_q(e, mq, raw);
  out[2 * q] = F.r;
  out[2 * q + 1] = F.i;
}

// ----- FP64 exact-triangle kernel -------------------------------------------------------------

// One flat triangle, src/ff/cpu/ff_tri_cpu.cpp:46-153 (phase exp(-i q.v)); mq is already
// negated for the exp(+i q.r) convention.
__device__ cd exact_triangle(const cd mq[3], double q_sqr, const double* __restrict__ v, double tiny) {
  double e[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    e[0][i] = v[3 + i] - v[i];
    e[1][i] = v[6 + i] - v[3 + i];
    e[2][i] = v[i] - v[6 + i];
  }
  double nt[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                  e[0][0] * e[1][1] - e[0][1] * e[1][0]};
  double ntn = sqrt(nt[0] * nt[0] + nt[1] * nt[1] + nt[2] * nt[2]);
  double area = 0.5 * ntn;
#pragma unroll
  for (int i = 0; i < 3; ++i) nt[i] /= ntn;
  cd q_nt = mq[0] * nt[0] + mq[1] * nt[1] + mq[2] * nt[2];
  double proj_tq = q_sqr - norm2(q_nt);
  if (fabs(proj_tq) < tiny) {
    cd q_v = mq[0] * v[0] + mq[1] * v[1] + mq[2] * v[2];
    return (area / q_sqr) * (mul_i(q_nt) * cexpi_d(-q_v));
  }
  cd ff = cd{0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double ne[3] = {e[k][1] * nt[2] - e[k][2] * nt[1], e[k][2] * nt[0] - e[k][0] * nt[2],
                    e[k][0] * nt[1] - e[k][1] * nt[0]};
    double nen = sqrt(ne[0] * ne[0] + ne[1] * ne[1] + ne[2] * ne[2]);
    cd q_ne = (mq[0] * ne[0] + mq[1] * ne[1] + mq[2] * ne[2]) * (1.0 / nen);
    double proj_eq = proj_tq - norm2(q_ne);
    double elen = sqrt(e[k][0] * e[k][0] + e[k][1] * e[k][1] + e[k][2] * e[k][2]);
    const double* va = v + 3 * k;
    const double* vb = v + 3 * ((k + 1) % 3);
    cd q_va = mq[0] * va[0] + mq[1] * va[1] + mq[2] * va[2];
    if (fabs(proj_eq) < tiny) {
      double f0 = elen / (q_sqr * proj_tq);
      ff = ff + f0 * (-(q_nt * q_ne) * cexpi_d(-q_va));
    } else {
      double f0 = q_sqr * proj_tq * proj_eq;
      cd q_nv = (mq[0] * e[k][0] + mq[1] * e[k][1] + mq[2] * e[k][2]) * (1.0 / elen);
      cd q_vb = mq[0] * vb[0] + mq[1] * vb[1] + mq[2] * vb[2];
      // +i: the reference's -i here (src/ff/cpu/ff_tri_cpu.cpp:132,146) has the opposite sign of
      // its cases 1-2 and of the volume integral
      cd c0 = mul_i(q_nt * q_ne * q_nv);
      ff = ff + (1.0 / f0) * (c0 * (cexpi_d(-q_va) - cexpi_d(-q_vb)));
    }
  }
  return ff;
}

__global__ __launch_bounds__(kWG) void k_ff_exact64(QView g, Rot R, double sign, const double* __restrict__ tri,
                                                    int64_t ntri, int64_t tchunk, int nchunks, double tiny,
                                                    double* __restrict__ part, double* __restrict__ out, Epi e) {
  int64_t q = (int64_t)blockIdx.x * kWG + threadIdx.x;
  const int chunk = blockIdx.y;
  const int64_t t0 = (int64_t)chunk * tchunk;
  const int64_t t1 = min(ntri, t0 + tchunk);
  cd mq[3];
  load_q(g, R, q < g.nq ? q : g.nq - 1, mq);
  cd ms[3] = {sign * mq[0], sign * mq[1], sign * mq[2]};
  double q2 = norm2(mq[0]) + norm2(mq[1]) + norm2(mq[2]);
  cd acc = cd{0.0, 0.0};
  if (q2 != 0.0)
    for (int64_t t = t0; t < t1; ++t) acc = acc + exact_triangle(ms, q2, tri + 9 * t, tiny);
  if (q >= g.nq) return;
  if (nchunks == 1) {
    cd F = q2 == 0.0 ? cd{e.volume, 0.0} : acc;
    if (e.has_tv) F = F * cexpi_d(mq[0] * e.tv.v[0] + mq[1] * e.tv.v[1] + mq[2] * e.tv.v[2]);
    out[2 * q] = F.r;
    out[2 * q + 1] = F.i;
  } else {
    part[2 * ((int64_t)chunk * g.nq + q)] = acc.r;
    part[2 * ((int64_t)chunk * g.nq + q) + 1] = acc.i;
  }
}

__global__ __launch_bounds__(kWG) void k_ff_exact_finish(QView g, Rot R, int nchunks, const double* __restrict__ part,
                                                         double* __restrict__ out, Epi e) {
  int64_t q = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (q >= g.nq) return;
  cd mq[3];
  load_q(g, R, q, mq);
  double q2 = norm2(mq[0]) + norm2(mq[1]) + norm2(mq[2]);
  cd F = cd{0.0, 0.0};
  for (int c = 0; c < nchunks; ++c) F = F + cd{part[2 * ((int64_t)c * g.nq + q)], part[2 * ((int64_t)c * g.nq + q) + 1]};
  if (q2 == 0.0) F = cd{e.volume, 0.0};
  if (e.has_tv) F = F * cexpi_d(mq[0] * e.tv.v[0] + mq[1] * e.tv.v[1] + mq[2] * e.tv.v[2]);
  out[2 * q] = F.r;
  out[2 * q + 1] = F.i;
}

// ----- FP32 compensated exact-triangle kernel -------------------------------------------------
//
// The flat-triangle integral is the centroid term of the approximate method times a factor:
//     int_T exp(i q.r) dS = s exp(i q.c) S(q),  S = -2 g[x1, x2, x3],  g(x) = exp(i x),
// the second divided difference of exp(i x) at x_j = q.d_j, d_j = v_j - c. Two evaluations, both
// free of the cancellations the closed form (an edge sum with 1/(q_t^2 q_e^2) factors) has near
// degenerate directions:
//   * series: S = 2 sum_n i^n h_n(x) / (n+2)!, h_n the complete homogeneous symmetric polynomial.
//     The x_j sum to zero, so h_n = p h_{n-2} + r h_{n-3} with p = x1^2 + x1 x2 + x2^2 and
//     r = -x1 x2 (x1 + x2). With a = max|q| (over the wave) x max|d| (over the 128-triangle block),
//     the FP32 truncation plus rounding error of S stays below 2e-7 (absolute; |S| <= 1) for
//     8/10/12/16/18 terms up to a = 0.5/1/2/3/4;
//   * divided differences, for blocks with a > 4: with x sorted lo <= md <= hi,
//     g[u, v] = i exp(i (u+v)/2) sinc((v-u)/2) (no cancellation at any spacing) and
//     S = -2 (g[md, hi] - g[lo, md]) / (hi - lo); a triangle (lane) whose spread hi - lo is below 1
//     takes the 10-term series instead (then max|x_j| <= 2/3). FP32 error below 3.2e-7 at spread 1,
//     falling as the spread grows.

// c_n i^n split: even n feed Re S with sign (-1)^(n/2), odd n feed Im S with sign (-1)^((n-1)/2)
__device__ constexpr float exact_coef(int n) {
  double f = 1.0;
  for (int j = 3; j <= n + 2; ++j) f *= j;
  return (float)(((n / 2) % 2 ? -2.0 : 2.0) / (2.0 * f));
}

// S for x1 = q.d1, x2 = q.d2 (radians) with NT terms (NT even, >= 4).
template <int NT>
__device__ inline void exact_series(float x1, float x2, float& Sr, float& Si) {
  const float p = __builtin_fmaf(x1, x1, x2 * (x1 + x2));
  const float r = -(x1 * x2) * (x1 + x2);
  // h0 = 1, h1 = 0, h2 = p, h3 = r
  Sr = __builtin_fmaf(exact_coef(2), p, 1.0f);
  Si = exact_coef(3) * r;
  float h1 = 0.0f, h2 = p, h3 = r;
#pragma unroll
  for (int n = 4; n < NT; n += 2) {
    const float he = __builtin_fmaf(p, h2, r * h1);  // h_n
    const float ho = __builtin_fmaf(p, h3, r * h2);  // h_{n+1}
    Sr = __builtin_fmaf(exact_coef(n), he, Sr);
    Si = __builtin_fmaf(exact_coef(n + 1), ho, Si);
    h1 = h3;
    h2 = he;
    h3 = ho;
  }
}

// sin(h)/h: Taylor to h^8 for h^2 < 1/4 (truncation below 3e-11), else the hardware sine (turns).
__device__ inline float sinc_small_safe(float h) {
  const float h2 = h * h;
  const float t = __builtin_fmaf(
      h2, __builtin_fmaf(h2, __builtin_fmaf(h2, __builtin_fmaf(h2, 1.0f / 362880.0f, -1.0f / 5040.0f), 1.0f / 120.0f),
                         -1.0f / 6.0f),
      1.0f);
  const float big = __builtin_amdgcn_sinf(h * (float)kInvTwoPi) * __builtin_amdgcn_rcpf(h);
  return h2 < 0.25f ? t : big;
}

// g[u, v] = i exp(i (u+v)/2) sinc((v-u)/2)
__device__ inline void exp_dd1(float u, float v, float& gr, float& gi) {
  const float m = 0.5f * (u + v) * (float)kInvTwoPi;  // turns
  const float sc = sinc_small_safe(0.5f * (v - u));
  gr = -__builtin_amdgcn_sinf(m) * sc;
  gi = __builtin_amdgcn_cosf(m) * sc;
}

// S by divided differences (spread >= 1) or the 10-term series.
__device__ inline void exact_dd(float x1, float x2, float& Sr, float& Si) {
  const float x3 = -(x1 + x2);
  const float lo = fminf(fminf(x1, x2), x3), hi = fmaxf(fmaxf(x1, x2), x3);
  const float md = __builtin_amdgcn_fmed3f(x1, x2, x3);
  if (hi - lo < 1.0f) {
    exact_series<10>(x1, x2, Sr, Si);
    return;
  }
  float ar, ai, br, bi;
  exp_dd1(md, hi, ar, ai);
  exp_dd1(lo, md, br, bi);
  const float w = -2.0f * __builtin_amdgcn_rcpf(hi - lo);
  Sr = (ar - br) * w;
  Si = (ai - bi) * w;
}

// One block of 128 triangles in FP32, block-relative: acc_k += sum_t (q_k.s n_t) exp(i q_k.(c_t - o_b)) S_t(q_k).
// NT: series terms, 0 for divided differences.
template <int QPT, int NT>
__device__ inline void exact32_block(const float* __restrict__ rb, const float* __restrict__ db, const float (&mx)[QPT],
                                     const float (&my)[QPT], const float (&mz)[QPT], const float (&ux)[QPT],
                                     const float (&uy)[QPT], const float (&uz)[QPT], float (&sr)[QPT],
                                     float (&si)[QPT]) {
#pragma unroll 2
  for (int t = 0; t < kBlockTri; ++t) {
    const float* r = rb + 8 * t;
    const float* d = db + 8 * t;
    const float snx = r[0], sny = r[1], snz = r[2], cx = r[3], cy = r[4], cz = r[5];
    const float d1x = d[0], d1y = d[1], d1z = d[2], d2x = d[3], d2y = d[4], d2z = d[5];
#pragma unroll
    for (int k = 0; k < QPT; ++k) {
      const float ph = ux[k] * cx + uy[k] * cy + uz[k] * cz;
      const float s = __builtin_amdgcn_sinf(ph);
      const float c = __builtin_amdgcn_cosf(ph);
      const float qn = mx[k] * snx + my[k] * sny + mz[k] * snz;
      const float x1 = mx[k] * d1x + my[k] * d1y + mz[k] * d1z;
      const float x2 = mx[k] * d2x + my[k] * d2y + mz[k] * d2z;
      float Sr, Si;
      if (NT == 0)
        exact_dd(x1, x2, Sr, Si);
      else
        exact_series<NT == 0 ? 4 : NT>(x1, x2, Sr, Si);
      const float er = qn * c, ei = qn * s;
      sr[k] = __builtin_fmaf(er, Sr, __builtin_fmaf(-ei, Si, sr[k]));
      si[k] = __builtin_fmaf(er, Si, __builtin_fmaf(ei, Sr, si[k]));
    }
  }
}

// Largest |q|^2 over the q-points of this thread's wave (the kWG-strided tile layout below), from
// the unrotated real grid (a rotation keeps |q|).
template <int QPT>
__device__ inline float wave_q2max(const QView& g, int64_t qbase) {
  float q2m = 0.0f;
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    int64_t q = qbase + (int64_t)k * kWG;
    q = q < g.nq ? q : g.nq - 1;
    const int64_t p = q % g.npix;
    const double qx = g.qx[p], qy = g.qy[p], qz = g.qz[2 * q];
    q2m = fmaxf(q2m, (float)(qx * qx + qy * qy + qz * qz));
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) q2m = fmaxf(q2m, __shfl_xor(q2m, off));
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(q2m)));
}

// Raw sums P = sum_t (q.s n_t) exp(i q.c_t) S_t(q) over blocks [b0, b1) for q = qsign * (R q_grid)
// (qsign -1 gives the reference's phase exp(-i q.r)); real q only. FP32 inside a block, FP64
// across blocks, like the approximate FP32C kernel.
template <int QPT>
__global__ __launch_bounds__(kWG) void k_ff_exact32(QView g, Rot R, double qsign, const float* __restrict__ rec,
                                                    const float* __restrict__ dv, const double* __restrict__ borig,
                                                    const float* __restrict__ bdmax, int64_t nblocks, int64_t bchunk,
                                                    int nchunks, double* __restrict__ part, double* __restrict__ out) {
  const int64_t qbase = (int64_t)blockIdx.x * (kWG * QPT) + threadIdx.x;
  const int chunk = blockIdx.y;
  const int64_t b0 = (int64_t)chunk * bchunk;
  const int64_t b1 = min(nblocks, b0 + bchunk);
  const float q2m = wave_q2max<QPT>(g, qbase);
  float mx[QPT], my[QPT], mz[QPT], ux[QPT], uy[QPT], uz[QPT];
  double ar[QPT], ai[QPT];
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    int64_t q = qbase + (int64_t)k * kWG;
    cd mq[3];
    load_q<false>(g, R, q < g.nq ? q : g.nq - 1, mq);
    const double qx = qsign * mq[0].r, qy = qsign * mq[1].r, qz = qsign * mq[2].r;
    mx[k] = (float)qx;
    my[k] = (float)qy;
    mz[k] = (float)qz;
    ux[k] = (float)(qx * kInvTwoPi);
    uy[k] = (float)(qy * kInvTwoPi);
    uz[k] = (float)(qz * kInvTwoPi);
    ar[k] = 0.0;
    ai[k] = 0.0;
  }
  for (int64_t b = b0; b < b1; ++b) {
    // a^2 = (max |q|)^2 (max |d|)^2: the evaluation is chosen per block for the whole wave
    const float bd = bdmax[b];
    const float a2 = q2m * (bd * bd);
    float sr[QPT], si[QPT];
#pragma unroll
    for (int k = 0; k < QPT; ++k) sr[k] = si[k] = 0.0f;
    const float* rb = rec + (int64_t)8 * kBlockTri * b;
    const float* db = dv + (int64_t)8 * kBlockTri * b;
    if (a2 <= 0.25f)
      exact32_block<QPT, 8>(rb, db, mx, my, mz, ux, uy, uz, sr, si);
    else if (a2 <= 1.0f)
      exact32_block<QPT, 10>(rb, db, mx, my, mz, ux, uy, uz, sr, si);
    else if (a2 <= 4.0f)
      exact32_block<QPT, 12>(rb, db, mx, my, mz, ux, uy, uz, sr, si);
    else if (a2 <= 9.0f)
      exact32_block<QPT, 16>(rb, db, mx, my, mz, ux, uy, uz, sr, si);
    else if (a2 <= 16.0f)
      exact32_block<QPT, 18>(rb, db, mx, my, mz, ux, uy, uz, sr, si);
    else
      exact32_block<QPT, 0>(rb, db, mx, my, mz, ux, uy, uz, sr, si);
    // rotate by exp(i q.o_b): FP64 phase in turns, reduced in FP64
    const double ox = borig[4 * b], oy = borig[4 * b + 1], oz = borig[4 * b + 2];
#pragma unroll
    for (int k = 0; k < QPT; ++k) {
      int64_t q = qbase + (int64_t)k * kWG;
      cd mq[3];
      load_q<false>(g, R, q < g.nq ? q : g.nq - 1, mq);
      const double tb = qsign * (mq[0].r * ox + mq[1].r * oy + mq[2].r * oz) * kInvTwoPi;
      const float fr = (float)(tb - rint(tb));
      const double cb = (double)__builtin_amdgcn_cosf(fr), sb = (double)__builtin_amdgcn_sinf(fr);
      ar[k] += (double)sr[k] * cb - (double)si[k] * sb;
      ai[k] += (double)sr[k] * sb + (double)si[k] * cb;
    }
  }
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    int64_t q = qbase + (int64_t)k * kWG;
    if (q >= g.nq) continue;
    double* dst = nchunks == 1 ? out + 2 * q : part + 2 * ((int64_t)chunk * g.nq + q);
    dst[0] = ar[k];
    dst[1] = ai[k];
  }
}

// Epilogue of k_ff_exact32: sums the chunks in order and finishes at q = qsign R q_grid (small-|q|
// series from the flat-triangle moments); the translation phase stays exp(i (R q_grid).t), as in
// the FP64 exact kernel.
__global__ __launch_bounds__(kWG) void k_ff_exact32_finish(QView g, Rot R, double qsign, int nchunks,
                                                           const double* __restrict__ part, double* __restrict__ out,
                                                           Epi e) {
  int64_t q = (int64_t)blockIdx.x * kWG + threadIdx.x;
  if (q >= g.nq) return;
  cd mq[3];
  load_q<false>(g, R, q, mq);
  double pr = 0.0, pi = 0.0;
  if (nchunks == 1) {
    pr = out[2 * q];
    pi = out[2 * q + 1];
  }
  for (int c = 0; nchunks > 1 && c < nchunks; ++c) {
    pr += part[2 * ((int64_t)c * g.nq + q)];
    pi += part[2 * ((int64_t)c * g.nq + q) + 1];
  }
  const cd qe[3] = {qsign * mq[0], qsign * mq[1], qsign * mq[2]};
  Epi e0 = e;
  e0.has_tv = 0;
  cd F = finish_q(e0, qe, cd{pr, pi});
  if (e.has_tv) F = F * cexpi_d(mq[0] * e.tv.v[0] + mq[1] * e.tv.v[1] + mq[2] * e.tv.v[2]);
  out[2 * q] = F.r;
  out[2 * q + 1] = F.i;
}

// ----- host-side launch planning --------------------------------------------------------------

struct Plan {
  int qpt;
  unsigned nqt;
  int nchunks;
  int64_t per_chunk;  // blocks (FP32C) or triangles (FP64 / exact) per chunk
};

// Enough workgroups to fill 256 CUs several times over, without splitting below one block, and
// at most max_chunks chunks.
static Plan make_plan(int64_t nq, int64_t units, int64_t min_units, int qpt, int target_wg,
                      int64_t max_chunks = INT64_MAX) {
  Plan p;
  p.qpt = qpt;
  p.nqt = ceil_div(nq, (int64_t)kWG * qpt);
  int64_t want = std::max<int64_t>(1, (target_wg + p.nqt - 1) / p.nqt);
  int64_t maxc = std::max<int64_t>(1, units / std::max<int64_t>(1, min_units));
  want = std::min(std::min(want, maxc), std::max<int64_t>(1, max_chunks));
  p.per_chunk = (units + want - 1) / want;
  p.nchunks = (int)((units + p.per_chunk - 1) / p.per_chunk);
  return p;
}

template <int QPT>
static void launch_approx64(hipStream_t s, const Plan& p, QView g, Rot R, const Mesh& m, bool cq, double* part,
                            double* out, Epi e) {
  dim3 grid(p.nqt, p.nchunks);
  if (cq)
    k_ff_approx64<QPT, true><<<grid, kWG, 0, s>>>(g, R, m.d_rec64, m.ntri, p.per_chunk, p.nchunks, part, out, e);
  else
    k_ff_approx64<QPT, false><<<grid, kWG, 0, s>>>(g, R, m.d_rec64, m.ntri, p.per_chunk, p.nchunks, part, out, e);
}

template <int QPT>
static void launch_approx32(hipStream_t s, const Plan& p, QView g, Rot R, const Mesh& m, bool cq, bool kahan,
                            double* part, double* out, Epi e) {
  dim3 grid(p.nqt, p.nchunks);
#define HGX_L32(CQ, KH) \
  k_ff_approx32<QPT, CQ, KH><<<grid, kWG, 0, s>>>(g, R, m.d_rec32, m.d_borig, m.nblocks, p.per_chunk, p.nchunks, part, out, e)
  if (cq) {
    if (kahan) HGX_L32(true, true); else HGX_L32(true, false);
  } else {
    if (kahan) HGX_L32(false, true); else HGX_L32(false, false);
  }
#undef HGX_L32
}

}  // namespace hgx

using namespace hgx;

// Tuning knobs for the numeric kernels (set by the autotuner; defaults are measured choices).
extern "C" int hgx_tune_numeric(hgx_ctx* h, int qpt64, int qpt32, int target_wg) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  auto ok = [](int v) { return v == 1 || v == 2 || v == 4 || v == 8; };
  if (!ok(qpt64) || !ok(qpt32) || target_wg < 1) return HGX_ERR_INVALID_ARG;
  c->tune.qpt64 = qpt64;
  c->tune.qpt32 = qpt32;
  c->tune.target_wg = target_wg;
  return HGX_OK;
}

namespace {

// Device memory of one numeric form-factor call (partial sums of triangle chunks, rotated records):
// the context's workspace, grown on demand (fixed == nullptr), or a fixed buffer of cap bytes that
// the call must fit in, choosing fewer chunks where needed (the allocation-free entry point).
struct Work {
  char* fixed = nullptr;
  size_t cap = 0;
};

int take_work(Context* c, const Work& w, size_t bytes, char** out) {
  if (!w.fixed) {
    if (int rc = ensure_work(c, bytes)) return rc;
    *out = static_cast<char*>(c->d_work);
    return HGX_OK;
  }
  if (bytes > w.cap)
    return set_error(c, HGX_ERR_INVALID_ARG, "ff_numeric_noalloc: needs %zu bytes of scratch, the maximum is %zu",
                     bytes, w.cap);
  *out = w.fixed;
  return HGX_OK;
}

// Fewest chunk count reductions: with a fixed buffer, at most as many chunks of nq complex partial
// sums as fit beside `reserved` bytes (one chunk needs none: it writes the output directly).
int64_t chunk_limit(const Work& w, int64_t nq, size_t reserved, size_t value_bytes = sizeof(double)) {
  if (!w.fixed) return INT64_MAX;
  if (reserved >= w.cap) return 1;
  const int64_t n = (int64_t)((w.cap - reserved) / ((size_t)nq * 2 * value_bytes));
  return n >= 2 ? n : 1;
}

// Launch timing (hgx_profile_begin) only for calls that may allocate: event creation is not done by
// the allocation-free entry point.
struct OptTimer {
  alignas(LaunchTimer) unsigned char buf[sizeof(LaunchTimer)];
  LaunchTimer* t = nullptr;
  OptTimer(Context* c, bool main, bool on) {
    if (on) t = new (buf) LaunchTimer(c, main);
  }
  ~OptTimer() {
    if (t) t->~LaunchTimer();
  }
};

// Series orders of the Ewald-grid matrix-core path for this call, or -1 when another kernel runs:
// the FP32C approximate method on a grid with the Ewald factorization that is not a product grid,
// with the matrix cores and block accumulation allowed, and few enough orders for the mesh's radius.
int ewald_path_orders(const Context* c, const Mesh& m, int method, int precision, int flags) {
  if (method != HGX_FF_APPROX || precision != HGX_PREC_FP32C || c->product || !c->ew.ok) return -1;
  if (flags & (HGX_FLAG_NO_MFMA | HGX_FLAG_ACC_KAHAN)) return -1;
  return ewald_orders(c, m.radius);
}

int ff_numeric_run(Context* c, int mesh_id, int method, int precision, int flags, const double rot[9],
                   const double transvec[3], double* ff, const Work& work) {
  const bool timed = work.fixed == nullptr;
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "ff_numeric: no q-grid");
  if (mesh_id < 0 || mesh_id >= (int)c->meshes.size() || !c->meshes[mesh_id].live)
    return set_error(c, HGX_ERR_BAD_HANDLE, "ff_numeric: unknown mesh %d", mesh_id);
  if (!ff) return set_error(c, HGX_ERR_INVALID_ARG, "ff_numeric: null output");
  if (precision != HGX_PREC_FP64 && precision != HGX_PREC_FP32C)
    return set_error(c, HGX_ERR_INVALID_ARG, "ff_numeric: bad precision %d", precision);
  const Mesh& m = c->meshes[mesh_id];
  if (int rc = set_device(c)) return rc;
  QView g{c->d_qx, c->d_qy, c->d_qz, c->npix, c->nq};
  Rot R = make_rot(rot);
  Epi e;
  e.tv = make_vec(transvec);
  e.has_tv = transvec != nullptr;
  e.volume = m.volume;
  e.moments = m.d_moments;
  for (int k = 0; k < 3; ++k) e.origin[k] = m.origin[k];
  e.radius = m.radius;
  e.use_series = 0;
  const bool cq = !c->qz_real;  // a rotation keeps real q real
  if (method == HGX_FF_APPROX && c->product && !(flags & HGX_FLAG_NO_MFMA)) {
    // product grid: complex matrix products on the matrix cores (product_mfma.hip; FP64:
    // product_fp64.hip)
    const bool f64 = precision == HGX_PREC_FP64;
    e.use_series = f64 ? 0 : 1;
    e.series_qr = kSeriesQRSplit;
    const bool identity = rot == nullptr || (R.m[0] == 1 && R.m[1] == 0 && R.m[2] == 0 && R.m[3] == 0 &&
                                             R.m[4] == 1 && R.m[5] == 0 && R.m[6] == 0 && R.m[7] == 0 && R.m[8] == 1);
    ProductItem it{};
    it.qx = c->d_qx;
    it.qy = c->d_qy;
    it.qz = c->d_qz;
    it.nblocks = m.nblocks;
    it.nq = c->nq;
    it.ncol = c->ncol;
    it.nz = c->nbranch * c->nrow;
    it.snmax = (float)m.snmax;
    it.rmax = (float)(std::sqrt(m.origin[0] * m.origin[0] + m.origin[1] * m.origin[1] + m.origin[2] * m.origin[2]) +
                      m.radius);
    it.ntri = m.ntri;
    const int pmode = f64 ? product_mode64(c) : product_mode(c);
    const int64_t tiles = (int64_t)ceil_div(it.nz, product_zrows(pmode)) * ceil_div(it.ncol, product_ycols(pmode));
    // enough waves to fill the chip several times over, without splitting below one block
    const int64_t upw = product_is_shared(pmode) ? 1 : kProductWaves;  // work units per workgroup
    int64_t want = std::max<int64_t>(1, (int64_t)(upw * c->tune.target_wg + tiles - 1) / tiles);
    want = std::min<int64_t>(want, m.nblocks);
    // chunks accumulate in FP32: at most product_blocks_per_unit(c) blocks each
    it.bchunk = std::min<int64_t>((m.nblocks + want - 1) / want, product_blocks_per_unit(c));
    size_t rec_bytes = identity ? 0
                       : f64    ? (size_t)m.ntri * 8 * sizeof(double)
                                : (size_t)m.nblocks * kBlockTri * 8 * sizeof(float) + (size_t)m.nblocks * 4 * sizeof(double);
    {
      // a fixed buffer may hold fewer chunks: longer chunks then (their FP32 sums span more blocks)
      const int64_t lim = chunk_limit(work, c->nq, rec_bytes, f64 ? sizeof(double) : sizeof(float));
      if ((m.nblocks + it.bchunk - 1) / it.bchunk > lim) it.bchunk = (m.nblocks + lim - 1) / lim;
    }
    it.nchunks = (int)((m.nblocks + it.bchunk - 1) / it.bchunk);
    if ((double)tiles * it.nchunks > 2.0e9) return set_error(c, HGX_ERR_INVALID_ARG, "ff_numeric: grid too large");
    // partial sums of the chunks: FP64 for the FP64 kernel, FP32 for the FP32C ones (their sums are FP32)
    const size_t pbytes = f64 ? sizeof(double) : sizeof(float);
    size_t part_bytes = it.nchunks > 1 ? (size_t)it.nchunks * c->nq * 2 * pbytes : 0;
    char* w = nullptr;
    if (part_bytes + rec_bytes > 0)
      if (int rc = take_work(c, work, part_bytes + rec_bytes, &w)) return rc;
    it.part = it.nchunks > 1 && f64 ? reinterpret_cast<double*>(w) : nullptr;
    it.part32 = it.nchunks > 1 && !f64 ? reinterpret_cast<float*>(w) : nullptr;
    it.out = ff;
    if (identity) {
      it.rec = m.d_rec32;
      it.borig = m.d_borig;
      it.rec64 = m.d_rec64;
    } else if (f64) {
      double* rr = reinterpret_cast<double*>(w + part_bytes);
      {
        OptTimer lt(c, false, timed);
        if (int rc = product_rotate_records64(c, m, R, rr, c->stream)) return rc;
      }
      it.rec64 = rr;
    } else {
      float* rr = reinterpret_cast<float*>(w + part_bytes);
      double* ro = reinterpret_cast<double*>(w + part_bytes + (size_t)m.nblocks * kBlockTri * 8 * sizeof(float));
      {
        OptTimer lt(c, false, timed);
        if (int rc = product_rotate_mesh(c, m, R, rr, ro, c->stream)) return rc;
      }
      it.rec = rr;
      it.borig = ro;
    }
    {
      OptTimer lt(c, true, timed);
      product_launch_single(c->stream, it, cq, pmode);
    }
    HGX_CHECK_LAUNCH(c);
    {
      OptTimer lt(c, false, timed);
      k_ff_finish<<<ceil_div(c->nq, kWG), kWG, 0, c->stream>>>(g, R, it.nchunks, it.part, it.part32, ff, e);
      HGX_CHECK_LAUNCH(c);
    }
    return HGX_OK;
  }
  const int ew_orders = ewald_path_orders(c, m, method, precision, flags);
  if (ew_orders >= 0) {
    // Ewald-sphere grid (QGrid::create, detector): matrix products with a short series in the
    // per-row scale of q_par (ewald_mfma.hip)
    e.use_series = 1;
    e.series_qr = kSeriesQRSplit;
    const bool identity = rot == nullptr || (R.m[0] == 1 && R.m[1] == 0 && R.m[2] == 0 && R.m[3] == 0 &&
                                             R.m[4] == 1 && R.m[5] == 0 && R.m[6] == 0 && R.m[7] == 0 && R.m[8] == 1);
    EwaldItem it{};
    ewald_fill_item(c, m, R, it);
    const int64_t tiles = (int64_t)it.ntile * ewald_col_groups(it.ncol);
    int64_t want = std::max<int64_t>(1, (int64_t)(c->tune.target_wg + tiles - 1) / tiles);
    want = std::min<int64_t>(want, m.nblocks);
    it.bchunk = std::min<int64_t>((m.nblocks + want - 1) / want, product_blocks_per_unit(c));
    const size_t rec_bytes =
        identity ? 0 : (size_t)m.nblocks * kBlockTri * 8 * sizeof(float) + (size_t)m.nblocks * 4 * sizeof(double);
    {
      const int64_t lim = chunk_limit(work, c->nq, rec_bytes, sizeof(float));
      if ((m.nblocks + it.bchunk - 1) / it.bchunk > lim) it.bchunk = (m.nblocks + lim - 1) / lim;
    }
    it.nchunks = (int)((m.nblocks + it.bchunk - 1) / it.bchunk);
    if ((double)tiles * it.nchunks > 2.0e9) return set_error(c, HGX_ERR_INVALID_ARG, "ff_numeric: grid too large");
    // FP32 partial sums of the chunks (the kernel's sums are FP32)
    const size_t part_bytes = it.nchunks > 1 ? (size_t)it.nchunks * c->nq * 2 * sizeof(float) : 0;
    char* w = nullptr;
    if (part_bytes + rec_bytes > 0)
      if (int rc = take_work(c, work, part_bytes + rec_bytes, &w)) return rc;
    it.part = nullptr;
    it.part32 = it.nchunks > 1 ? reinterpret_cast<float*>(w) : nullptr;
    it.out = ff;
    if (identity) {
      it.rec = m.d_rec32;
      it.borig = m.d_borig;
    } else {
      float* rr = reinterpret_cast<float*>(w + part_bytes);
      double* ro = reinterpret_cast<double*>(w + part_bytes + (size_t)m.nblocks * kBlockTri * 8 * sizeof(float));
      {
        OptTimer lt(c, false, timed);
        if (int rc = product_rotate_mesh(c, m, R, rr, ro, c->stream)) return rc;
      }
      it.rec = rr;
      it.borig = ro;
    }
    {
      OptTimer lt(c, true, timed);
      if (int rc = ewald_launch(c, it, cq, ew_orders)) return rc;
    }
    {
      OptTimer lt(c, false, timed);
      k_ff_finish<<<ceil_div(c->nq, kWG), kWG, 0, c->stream>>>(g, R, it.nchunks, it.part, it.part32, ff, e);
      HGX_CHECK_LAUNCH(c);
    }
    return HGX_OK;
  }
  if (method == HGX_FF_APPROX) {
    Plan p;
    if (precision == HGX_PREC_FP64)
      p = make_plan(c->nq, m.ntri, kBlockTri, c->tune.qpt64, c->tune.target_wg, chunk_limit(work, c->nq, 0));
    else
      p = make_plan(c->nq, m.nblocks, 1, c->tune.qpt32, c->tune.target_wg, chunk_limit(work, c->nq, 0));
    double* part = nullptr;
    if (p.nchunks > 1) {
      char* w = nullptr;
      if (int rc = take_work(c, work, (size_t)p.nchunks * c->nq * 2 * sizeof(double), &w)) return rc;
      part = reinterpret_cast<double*>(w);
    }
    {
    OptTimer lt(c, true, timed);
    if (precision == HGX_PREC_FP64) {
      switch (p.qpt) {
        case 1: launch_approx64<1>(c->stream, p, g, R, m, cq, part, ff, e); break;
        case 2: launch_approx64<2>(c->stream, p, g, R, m, cq, part, ff, e); break;
        case 4: launch_approx64<4>(c->stream, p, g, R, m, cq, part, ff, e); break;
        default: launch_approx64<8>(c->stream, p, g, R, m, cq, part, ff, e); break;
      }
    } else {
      e.use_series = 1;
      bool kahan = (flags & HGX_FLAG_ACC_KAHAN) != 0;
      switch (p.qpt) {
        case 1: launch_approx32<1>(c->stream, p, g, R, m, cq, kahan, part, ff, e); break;
        case 2: launch_approx32<2>(c->stream, p, g, R, m, cq, kahan, part, ff, e); break;
        case 4: launch_approx32<4>(c->stream, p, g, R, m, cq, kahan, part, ff, e); break;
        default: launch_approx32<8>(c->stream, p, g, R, m, cq, kahan, part, ff, e); break;
      }
    }
    }
    HGX_CHECK_LAUNCH(c);
    {
      OptTimer lt(c, false, timed);
      k_ff_finish<<<ceil_div(c->nq, kWG), kWG, 0, c->stream>>>(g, R, p.nchunks, part, nullptr, ff, e);
      HGX_CHECK_LAUNCH(c);
    }
    return HGX_OK;
  }
  if (method == HGX_FF_EXACT) {
    if (m.kind != HGX_MESH_EXACT9 || !m.d_tri9)
      return set_error(c, HGX_ERR_INVALID_ARG, "ff_numeric: exact method needs a vertex (exact9) mesh");
    double sign = (flags & HGX_FLAG_EXACT_REFERENCE_SIGN) ? 1.0 : -1.0;
    // degenerate-case threshold on q.n and q.e: the reference GPU path's 1e-13 (ff_tri_gpu.cu:80,100)
    // by default; its CPU path's DBL_MIN (ff_tri_cpu.cpp:87,107) on request. With DBL_MIN a case
    // decision can flip on the last bit of a nearly degenerate projection, and the general case then
    // divides by it: on a finely meshed solid most results become rounding noise.
    const double tiny = (flags & HGX_FLAG_EXACT_CPU_THRESHOLD) ? 2.2250738585072014e-308 : 1e-13;
    if (precision == HGX_PREC_FP32C) {
      if (cq)
        return set_error(c, HGX_ERR_UNSUPPORTED,
                         "ff_numeric: the FP32C exact method takes a real q-grid (FP64 handles complex qz)");
      if (!m.d_dv32 || !m.d_bdmax || !m.d_xmoments)
        return set_error(c, HGX_ERR_INVALID_ARG, "ff_numeric: mesh has no exact-triangle tables");
      const int qpt = c->tune.qpt32 >= 4 ? 4 : c->tune.qpt32 >= 2 ? 2 : 1;
      Plan p = make_plan(c->nq, m.nblocks, 1, qpt, c->tune.target_wg, chunk_limit(work, c->nq, 0));
      double* part = nullptr;
      if (p.nchunks > 1) {
        char* w = nullptr;
        if (int rc = take_work(c, work, (size_t)p.nchunks * c->nq * 2 * sizeof(double), &w)) return rc;
        part = reinterpret_cast<double*>(w);
      }
      e.moments = m.d_xmoments;
      e.radius = m.xradius;
      e.use_series = 1;
      const double qsign = -sign;  // phase exp(i qsign q.r)
      dim3 grid(p.nqt, p.nchunks);
      {
        OptTimer lt(c, true, timed);
#define HGX_LX32(Q)                                                                                      \
  k_ff_exact32<Q><<<grid, kWG, 0, c->stream>>>(g, R, qsign, m.d_rec32, m.d_dv32, m.d_borig, m.d_bdmax, m.nblocks, \
                                               p.per_chunk, p.nchunks, part, ff)
        if (p.qpt == 4) HGX_LX32(4);
        else if (p.qpt == 2) HGX_LX32(2);
        else HGX_LX32(1);
#undef HGX_LX32
      }
      HGX_CHECK_LAUNCH(c);
      {
        OptTimer lt(c, false, timed);
        k_ff_exact32_finish<<<ceil_div(c->nq, kWG), kWG, 0, c->stream>>>(g, R, qsign, p.nchunks, part, ff, e);
        HGX_CHECK_LAUNCH(c);
      }
      return HGX_OK;
    }
    if (precision != HGX_PREC_FP64)
      return set_error(c, HGX_ERR_UNSUPPORTED, "ff_numeric: bad precision for the exact method");
    Plan p = make_plan(c->nq, m.ntri, 64, 1, c->tune.target_wg, chunk_limit(work, c->nq, 0));
    double* part = nullptr;
    if (p.nchunks > 1) {
      char* w = nullptr;
      if (int rc = take_work(c, work, (size_t)p.nchunks * c->nq * 2 * sizeof(double), &w)) return rc;
      part = reinterpret_cast<double*>(w);
    }
    dim3 grid(p.nqt, p.nchunks);
    {
      OptTimer lt(c, true, timed);
      k_ff_exact64<<<grid, kWG, 0, c->stream>>>(g, R, sign, m.d_tri9, m.ntri, p.per_chunk, p.nchunks, tiny, part, ff, e);
    }
    HGX_CHECK_LAUNCH(c);
    if (p.nchunks > 1) {
      OptTimer lt(c, false, timed);
      k_ff_exact_finish<<<ceil_div(c->nq, kWG), kWG, 0, c->stream>>>(g, R, p.nchunks, part, ff, e);
      HGX_CHECK_LAUNCH(c);
    }
    return HGX_OK;
  }
  return set_error(c, HGX_ERR_INVALID_ARG, "ff_numeric: bad method %d", method);
}

}  // namespace

extern "C" int hgx_ff_numeric(hgx_ctx* h, int mesh_id, int method, int precision, int flags, const double rot[9],
                              const double transvec[3], double* ff) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  return ff_numeric_run(c, mesh_id, method, precision, flags, rot, transvec, ff, Work{});
}

extern "C" int hgx_ff_numeric_noalloc(hgx_ctx* h, int mesh_id, int method, int precision, int flags,
                                      const double rot[9], const double transvec[3], double* ff) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!c->d_scratch)
    return set_error(c, HGX_ERR_DEVICE, "ff_numeric_noalloc: the scratch array of the device was not resolved");
  return ff_numeric_run(c, mesh_id, method, precision, flags, rot, transvec, ff, Work{c->d_scratch, kScratchBytes});
}

extern "C" int hgx_ff_numeric_path(const hgx_ctx* h, int mesh_id, int method, int precision, int flags, int* path,
                                   int* orders) {
  const Context* c = reinterpret_cast<const Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!path) return HGX_ERR_INVALID_ARG;
  if (!c->d_qz) return HGX_ERR_NO_QGRID;
  if (mesh_id < 0 || mesh_id >= (int)c->meshes.size() || !c->meshes[mesh_id].live) return HGX_ERR_BAD_HANDLE;
  if (method != HGX_FF_APPROX && method != HGX_FF_EXACT) return HGX_ERR_INVALID_ARG;
  if (precision != HGX_PREC_FP64 && precision != HGX_PREC_FP32C) return HGX_ERR_INVALID_ARG;
  const int k = ewald_path_orders(c, c->meshes[mesh_id], method, precision, flags);
  if (orders) *orders = k;
  if (method == HGX_FF_APPROX && c->product && !(flags & HGX_FLAG_NO_MFMA)) *path = HGX_PATH_PRODUCT;
  else *path = k >= 0 ? HGX_PATH_EWALD : HGX_PATH_PAIR;
  return HGX_OK;
}

extern "C" int hgx_qgrid_is_ewald(const hgx_ctx* h, int* is_ewald, double* dv) {
  const Context* c = reinterpret_cast<const Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (is_ewald) *is_ewald = c->ew.ok ? 1 : 0;
  if (dv) *dv = c->ew.ok ? c->ew.dmax * c->ew.vmax : 0.0;
  return HGX_OK;
}
