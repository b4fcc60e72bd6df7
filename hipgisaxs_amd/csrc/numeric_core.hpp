// numeric_core.hpp -- pieces of the numeric (triangulated) form factor shared by the single-shape
// kernels (ff_numeric.hip) and the batched kernels (batch.hip): q-point loading, the per-q epilogue
// with the small-|q| moment series, and the FP32 block loop of the FP32C path.
#pragma once

#include "hgx_internal.hpp"

namespace hgx {

// Raw sum of one q-point over one chunk of triangles: in place (FP64) when the item has one chunk, else
// into the chunk's slot of the partial sums, FP32 (part32: the matrix-core kernels' FP32 sums, half the
// traffic) or FP64 (part).
__device__ inline void store_raw(int nchunks, double* out, double* part, float* part32, int64_t slot, int64_t q,
                                 double re, double im) {
  if (nchunks == 1) {
    out[2 * q] = re;
    out[2 * q + 1] = im;
  } else if (part32) {
    *reinterpret_cast<float2*>(part32 + 2 * slot) = make_float2((float)re, (float)im);
  } else {
    part[2 * slot] = re;
    part[2 * slot + 1] = im;
  }
}

// Sum over chunks of one q-point's partial sums, in chunk order, in FP64. Four chunks' loads are issued
// before their additions (one thread per q-point has little else to hide memory latency with).
__device__ inline cd sum_chunks(int nchunks, const double* part, const float* part32, int64_t nq, int64_t q) {
  double pr = 0.0, pi = 0.0;
  int c = 0;
  if (part32) {
    const float2* p = reinterpret_cast<const float2*>(part32) + q;
    for (; c + 4 <= nchunks; c += 4) {
      const float2 v0 = p[(int64_t)c * nq], v1 = p[(int64_t)(c + 1) * nq];
      const float2 v2 = p[(int64_t)(c + 2) * nq], v3 = p[(int64_t)(c + 3) * nq];
      pr += (double)v0.x;
      pi += (double)v0.y;
      pr += (double)v1.x;
      pi += (double)v1.y;
      pr += (double)v2.x;
      pi += (double)v2.y;
      pr += (double)v3.x;
      pi += (double)v3.y;
    }
    for (; c < nchunks; ++c) {
      const float2 v = p[(int64_t)c * nq];
      pr += (double)v.x;
      pi += (double)v.y;
    }
  } else {
    const double2* p = reinterpret_cast<const double2*>(part) + q;
    for (; c + 4 <= nchunks; c += 4) {
      const double2 v0 = p[(int64_t)c * nq], v1 = p[(int64_t)(c + 1) * nq];
      const double2 v2 = p[(int64_t)(c + 2) * nq], v3 = p[(int64_t)(c + 3) * nq];
      pr += v0.x;
      pi += v0.y;
      pr += v1.x;
      pi += v1.y;
      pr += v2.x;
      pi += v2.y;
      pr += v3.x;
      pi += v3.y;
    }
    for (; c < nchunks; ++c) {
      const double2 v = p[(int64_t)c * nq];
      pr += v.x;
      pi += v.y;
    }
  }
  return cd{pr, pi};
}

constexpr int kWG = 256;
constexpr double kSeriesQR = 0.1;
// Series threshold of the split-fp16 matrix-core paths: their operands carry about 22 bits, so the
// direct sum's cancellation near q = 0 (a loss of about 3 / (|q| R)) is kept further away; the order-8
// series is still good to 5e-10 of F at |q| R = 0.3.
constexpr double kSeriesQRSplit = 0.3;
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kInvTwoPi = 0.15915494309189533576888376337251;

struct QView {
  const double* __restrict__ qx;
  const double* __restrict__ qy;
  const double* __restrict__ qz;
  int64_t npix, nq;
};

struct Epi {
  Vec3 tv;
  int has_tv;
  double volume;
  // FP32C small-q series
  const double* moments;
  double origin[3];
  double radius;
  int use_series;
  double series_qr = kSeriesQR;  // the series replaces the direct sum below |q| R = series_qr
};

// Loads and rotates q-point q; with CQ = false the grid is known to be real and the imaginary
// parts are set to zero so the compiler drops them from the hot loops.
template <bool CQ = true>
__device__ inline void load_q(const QView& g, const Rot& R, int64_t q, cd mq[3]) {
  int64_t p = q % g.npix;
  cd qz = cd{g.qz[2 * q], CQ ? g.qz[2 * q + 1] : 0.0};
  rotate_q(R, g.qx[p], g.qy[p], qz, mq);
  if (!CQ) mq[0].i = mq[1].i = mq[2].i = 0.0;
}

// sum_{k=0..K} (i^k / k!) sum_t s_t (q.n_t) (q.(c_t - o))^k from the mesh moments. The moments are
// stored per order k, then per component i of n, then for a = k..0, b = k - a..0 (c = k - a - b):
// M_i[k; a, b] = sum_t s_t n_ti (c - o)_x^a (c - o)_y^b (c - o)_z^c. With the multinomial expansion
// of (q.(c - o))^k the sum is
//     sum_{a + b + c <= K} i^{a+b+c} (x^a / a!) (y^b / b!) (z^c / c!) sum_i q_i M_i[a + b + c; a, b],
// (x, y, z) = q, evaluated in one pass over the monomials: the three scaled powers are updated by one
// multiplication each, the moment index has a closed form, and loop counters are wave-uniform. This
// runs only for the few q-points with |q| R below the series threshold, but for the whole wave that
// holds one, so it is kept short.
__device__ inline cd moment_series(const double* __restrict__ M, const cd mq[3]) {
  static constexpr double kInv[kSeriesOrder + 1] = {1.0,       1.0,       1.0 / 2.0, 1.0 / 3.0, 1.0 / 4.0,
                                                    1.0 / 5.0, 1.0 / 6.0, 1.0 / 7.0, 1.0 / 8.0};
  static_assert(kSeriesOrder == 8, "table of 1/n");
  cd total = cd{0.0, 0.0};
  cd xa = cd{1.0, 0.0};
  for (int a = 0; a <= kSeriesOrder; ++a) {
    if (a > 0) xa = kInv[a] * (xa * mq[0]);
    cd yb = cd{1.0, 0.0};
    for (int b = 0; a + b <= kSeriesOrder; ++b) {
      if (b > 0) yb = kInv[b] * (yb * mq[1]);
      const cd xy = xa * yb;
      cd zc = cd{1.0, 0.0};
      for (int c = 0; a + b + c <= kSeriesOrder; ++c) {
        if (c > 0) zc = kInv[c] * (zc * mq[2]);
        const int k = a + b + c, A = k - a;
        const int nk = (k + 1) * (k + 2) / 2;                            // monomials of order k
        const int idx = k * (k + 1) * (k + 2) / 2 + A * (A + 1) / 2 + (A - b);  // component 0
        const cd w = M[idx] * mq[0] + M[idx + nk] * mq[1] + M[idx + 2 * nk] * mq[2];
        cd t = (xy * zc) * w;
        if (k & 1) t = mul_i(t);
        if (k & 2) t = -t;
        total = total + t;
      }
    }
  }
  return total;
}

// Final value of one q-point from the raw sum P = sum_t (q . s n) exp(i q . c) (FP64 path), or
// P = sum_b exp(i q.o_b) sum_{t in b} ... (FP32C path, identical value).
__device__ inline cd finish_q(const Epi& e, const cd mq[3], cd P) {
  double q2 = norm2(mq[0]) + norm2(mq[1]) + norm2(mq[2]);
  cd F;
  if (q2 == 0.0) {
    F = cd{e.volume, 0.0};
  } else if (e.use_series && sqrt(q2) * e.radius < e.series_qr) {
    cd S = moment_series(e.moments, mq);
    cd ph = mq[0] * e.origin[0] + mq[1] * e.origin[1] + mq[2] * e.origin[2];
    F = (1.0 / q2) * mul_mi(S * cexpi_d(ph));
  } else {
    F = (1.0 / q2) * mul_mi(P);
  }
  if (e.has_tv) F = F * cexpi_d(mq[0] * e.tv.v[0] + mq[1] * e.tv.v[1] + mq[2] * e.tv.v[2]);
  return F;
}

template <bool KAHAN>
struct Acc32 {
  float sr, si, cr, ci;
  __device__ inline void zero() { sr = si = cr = ci = 0.0f; }
  __device__ inline void add(float xr, float xi) {
    if (KAHAN) {
      float yr = xr - cr, yi = xi - ci;
      float tr = sr + yr, ti = si + yi;
      cr = (tr - sr) - yr;
      ci = (ti - si) - yi;
      sr = tr;
      si = ti;
    } else {
      sr += xr;
      si += xi;
    }
  }
};

// Sums blocks [b0, b1) of a mesh for QPT q-points q_k = qbase + k * kWG:
// P_k = sum_b exp(i q_k.o_b) sum_{t in b} (q_k . s n_t) exp(i q_k.(c_t - o_b)), FP32 inside a block
// (plain or Kahan), FP64 across blocks. Only FP32 copies of the rotated q stay live through the
// triangle loop; the FP64 q needed for the block phase is reloaded and re-rotated per block (a few
// cached loads per 128 triangles), which keeps the register count, and so the occupancy, up.
template <int QPT, bool CQ, bool KAHAN>
__device__ inline void approx32_blocks(const QView& g, const Rot& R, int64_t qbase, const float* __restrict__ rec,
                                       const double* __restrict__ borig, int64_t b0, int64_t b1, double (&ar)[QPT],
                                       double (&ai)[QPT]) {
  float mx[QPT], my[QPT], mz[QPT], ux[QPT], uy[QPT], uz[QPT];
  float mxi[QPT], myi[QPT], mzi[QPT];  // imaginary parts (complex q only)
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    int64_t q = qbase + (int64_t)k * kWG;
    cd mq[3];
    load_q<CQ>(g, R, q < g.nq ? q : g.nq - 1, mq);
    mx[k] = (float)mq[0].r;
    my[k] = (float)mq[1].r;
    mz[k] = (float)mq[2].r;
    ux[k] = (float)(mq[0].r * kInvTwoPi);
    uy[k] = (float)(mq[1].r * kInvTwoPi);
    uz[k] = (float)(mq[2].r * kInvTwoPi);
    // attenuation exp(-Im(q).c) = exp2(-log2(e) Im(q).c)
    mxi[k] = (float)(mq[0].i * 1.4426950408889634);
    myi[k] = (float)(mq[1].i * 1.4426950408889634);
    mzi[k] = (float)(mq[2].i * 1.4426950408889634);
    ar[k] = 0.0;
    ai[k] = 0.0;
  }
  for (int64_t b = b0; b < b1; ++b) {
    Acc32<KAHAN> acc[QPT];
#pragma unroll
    for (int k = 0; k < QPT; ++k) acc[k].zero();
    const float* rb = rec + (int64_t)8 * kBlockTri * b;
#pragma unroll 4
    for (int t = 0; t < kBlockTri; ++t) {
      const float* r = rb + 8 * t;
      const float snx = r[0], sny = r[1], snz = r[2], cx = r[3], cy = r[4], cz = r[5];
#pragma unroll
      for (int k = 0; k < QPT; ++k) {
        // v_sin/v_cos take turns and reduce internally (measured max abs error 1.25e-7 up to
        // +-250 turns, 2.6e-7 after an explicit v_fract); block-relative |ph| stays far below that
        float ph = ux[k] * cx + uy[k] * cy + uz[k] * cz;
        float s = __builtin_amdgcn_sinf(ph);
        float c = __builtin_amdgcn_cosf(ph);
        if (!CQ) {
          float qn = mx[k] * snx + my[k] * sny + mz[k] * snz;
          acc[k].add(qn * c, qn * s);
        } else {
          float att = __builtin_amdgcn_exp2f(-(mxi[k] * cx + myi[k] * cy + mzi[k] * cz));
          float qnr = (mx[k] * snx + my[k] * sny + mz[k] * snz) * att;
          float qni = (mxi[k] * snx + myi[k] * sny + mzi[k] * snz) * (att * 0.69314718055994531f);
          acc[k].add(qnr * c - qni * s, qnr * s + qni * c);
        }
      }
    }
    // rotate the block sum by exp(i q . o_b): FP64 phase in turns, reduced in FP64, then the
    // hardware sin/cos on the reduced FP32 angle
    const double ox = borig[4 * b], oy = borig[4 * b + 1], oz = borig[4 * b + 2];
#pragma unroll
    for (int k = 0; k < QPT; ++k) {
      int64_t q = qbase + (int64_t)k * kWG;
      cd mq[3];
      load_q<CQ>(g, R, q < g.nq ? q : g.nq - 1, mq);
      double tb = (mq[0].r * ox + mq[1].r * oy + mq[2].r * oz) * kInvTwoPi;
      float fr = (float)(tb - rint(tb));
      double cb = (double)__builtin_amdgcn_cosf(fr), sb = (double)__builtin_amdgcn_sinf(fr);
      if (CQ) {
        double m = exp(-(mq[0].i * ox + mq[1].i * oy + mq[2].i * oz));
        cb *= m;
        sb *= m;
      }
      double pr = (double)acc[k].sr, pi = (double)acc[k].si;
      if (KAHAN) {
        pr -= (double)acc[k].cr;
        pi -= (double)acc[k].ci;
      }
      ar[k] += pr * cb - pi * sb;
      ai[k] += pr * sb + pi * cb;
    }
  }
}

// Plain FP64 sum over triangles [t0, t1) (absolute coordinates, 8 doubles per record) for QPT
// q-points: P_k = sum_t (q_k . s n_t) exp(i q_k . c_t).
template <int QPT, bool CQ>
__device__ inline void approx64_tris(const cd (&mq)[QPT][3], const double* __restrict__ rec, int64_t t0, int64_t t1,
                                     double (&ar)[QPT], double (&ai)[QPT]) {
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    ar[k] = 0.0;
    ai[k] = 0.0;
  }
  for (int64_t t = t0; t < t1; ++t) {
    const double* r = rec + 8 * t;
    const double snx = r[0], sny = r[1], snz = r[2], cx = r[3], cy = r[4], cz = r[5];
#pragma unroll
    for (int k = 0; k < QPT; ++k) {
      if (!CQ) {
        double qn = mq[k][0].r * snx + mq[k][1].r * sny + mq[k][2].r * snz;
        double qt = mq[k][0].r * cx + mq[k][1].r * cy + mq[k][2].r * cz;
        double s, c;
        sincos(qt, &s, &c);
        ar[k] += qn * c;
        ai[k] += qn * s;
      } else {
        cd qn = mq[k][0] * snx + mq[k][1] * sny + mq[k][2] * snz;
        cd qt = mq[k][0] * cx + mq[k][1] * cy + mq[k][2] * cz;
        cd E = cexpi_d(qt);
        ar[k] += qn.r * E.r - qn.i * E.i;
        ai[k] += qn.r * E.i + qn.i * E.r;
      }
    }
  }
}

}  // namespace hgx
