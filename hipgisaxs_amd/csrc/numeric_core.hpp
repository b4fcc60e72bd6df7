// numeric_core.hpp -- pieces of the numeric (triangulated) form factor shared by the single-shape
// kernels (ff_numeric.hip) and the batched kernels (batch.hip): q-point loading, the per-q epilogue
// with the small-|q| moment series, and the FP32 block loop of the FP32C path.
#pragma once

#include "hgx_internal.hpp"

namespace hgx {

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

// sum_{k=0..K} (i^k / k!) sum_t s_t (q.n_t) (q.(c_t - o))^k from the mesh moments. Monomials are
// built by repeated multiplication (no indexed arrays, so nothing spills to scratch); this runs
// only for the few q-points with |q| R below the series threshold.
__device__ inline cd cpow_small(cd z, int n) {
  cd r = cd{1.0, 0.0};
  for (int i = 0; i < n; ++i) r = r * z;
  return r;
}

__device__ inline double fact_small(int n) {
  double f = 1.0;
  for (int j = 2; j <= n; ++j) f *= j;
  return f;
}

__device__ inline cd moment_series(const double* __restrict__ M, const cd mq[3]) {
  cd total = cd{0.0, 0.0};
  cd ik = cd{1.0, 0.0};
  int idx = 0;
  for (int k = 0; k <= kSeriesOrder; ++k) {
    const double fk = fact_small(k);
    cd Tk = cd{0.0, 0.0};
    for (int i = 0; i < 3; ++i) {
      cd s = cd{0.0, 0.0};
      for (int a = k; a >= 0; --a) {
        const cd xa = cpow_small(mq[0], a);
        for (int b = k - a; b >= 0; --b) {
          const int c = k - a - b;
          const double coef = fk / (fact_small(a) * fact_small(b) * fact_small(c));  // k!/(a! b! c!)
          s = s + (coef * M[idx++]) * (xa * cpow_small(mq[1], b) * cpow_small(mq[2], c));
        }
      }
      Tk = Tk + mq[i] * s;
    }
    total = total + (1.0 / fk) * (ik * Tk);
    ik = mul_i(ik);
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
