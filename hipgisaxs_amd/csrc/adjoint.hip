// adjoint.hip -- record adjoint of the approximate-triangle form factor.
//
// For L = Re sum_q W(q) F(q), with F(q) = -i / |q|^2 sum_t (q . A_t) exp(i q . c_t) (A_t = s_t n_t,
// src/ff/cpu/ff_tri_cpu.cpp:200-246) and W given per q-point, the gradient with respect to every
// triangle record is
//     dL/dA_t = Re sum_q Wt(q) q exp(i q.c_t),              Wt = -i W / |q|^2
//     dL/dc_t = Re i sum_q Wt(q) (q . A_t) q exp(i q.c_t) = Re i (M_t A_t),
//     M_t = sum_q Wt(q) q q^T exp(i q.c_t)
// plus, at q = 0 where F = V = sum_t A_t . c_t / 3, Re W(0) (c_t, A_t) / 3. Any mesh
// parameterisation's gradient follows by the chain rule through (A_t, c_t). The reference has no
// such path; it differentiates by finite differences (src/analyzer/hipgisaxs_fit_lmvm.cpp:190-224).
//
// Rotation: as in the forward kernels, records are rotated by R^T and the grid is left alone; the
// gradient is rotated back by R at the end.
//
// Plain FP64 (any grid): one thread per triangle, q-points in chunks (wave-uniform q data through
// scalar loads), FP64 partials per chunk reduced in order.
//
// FP32C on any other grid: the same loop in FP32 on block-relative records, the block's phase folded
// into the weights staged in LDS (k_adjoint32).
//
// FP32C on product grids q = (qx0, qy_y, qz_z): the sums are bilinear forms
//     B[f, g](t) = sum_{z, y} f(z) g(y) Wt(z, y) Ez_t(z) Ey_t(y)
// with f in {1, qz, qz^2} and g in {1, qy, qy^2}; V_t and M_t are combinations of the six forms
// B11, B1y, B1yy, Bz1, Bzy, Bzz. For each f the kernel forms X = (f Ez)^T Wt -- a complex matrix
// product (triangles x z) . (z x y) on v_mfma_f32_32x32x2_f32, Gauss's three real products -- and then
// contracts each row of X with g(y) Ey_t(y) across the wave's 32 columns. Phases come from FP64
// absolute centroids reduced to turns per table entry, so no block origins are needed.
#include "adjoint_core.hpp"
#include "numeric_core.hpp"

namespace hgx {

typedef float f32x16a __attribute__((ext_vector_type(16)));

constexpr int kAdjTile = 32;
constexpr int kAdjForms = 6;  // B11, B1y, B1yy, Bz1, Bzy, Bzz

// Rotated FP64 records (A' = R^T A, c' = R^T c) from the mesh's [ntri][8] FP64 records.
__global__ __launch_bounds__(256) void k_adj_rotate(const double* __restrict__ rec, int64_t ntri, Rot R,
                                                    double* __restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  const double* a = rec + 8 * t;
  double* o = out + 8 * t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o[k] = R.m[k] * a[0] + R.m[3 + k] * a[1] + R.m[6 + k] * a[2];
    o[3 + k] = R.m[k] * a[3] + R.m[3 + k] * a[4] + R.m[6 + k] * a[5];
  }
  o[6] = o[7] = 0.0;
}

// Wt = -i W / |q|^2 (0 at q = 0), FP64 or FP32.
template <typename T>
__global__ __launch_bounds__(256) void k_adj_weights(QView g, const double* __restrict__ W, T* __restrict__ wt) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.nq) return;
  int64_t p = q % g.npix;
  double qx = g.qx[p], qy = g.qy[p], zr = g.qz[2 * q], zi = g.qz[2 * q + 1];
  double q2 = qx * qx + qy * qy + zr * zr + zi * zi;
  double s = q2 == 0.0 ? 0.0 : 1.0 / q2;
  // -i (a + ib) = b - ia
  wt[2 * q] = (T)(W[2 * q + 1] * s);
  wt[2 * q + 1] = (T)(-W[2 * q] * s);
}

// w0 = sum over q-points with |q| = 0 of Re W(q), in a fixed order (deterministic): workgroup b sums
// the q-points [b per, (b + 1) per) by a fixed-order tree into parts[b]; k_adj_q0_sum adds the parts.
constexpr int kAdjQ0Parts = 512;

__device__ inline double block_sum256(double v) {
  __shared__ double s[256];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  return s[0];
}

__global__ __launch_bounds__(256) void k_adj_q0(QView g, const double* __restrict__ W, int64_t per,
                                                double* __restrict__ parts) {
  const int64_t q0 = (int64_t)blockIdx.x * per, q1 = min(g.nq, q0 + per);
  double acc = 0.0;
  for (int64_t q = q0 + threadIdx.x; q < q1; q += 256) {
    int64_t p = q % g.npix;
    double qx = g.qx[p], qy = g.qy[p], zr = g.qz[2 * q], zi = g.qz[2 * q + 1];
    if (qx * qx + qy * qy + zr * zr + zi * zi == 0.0) acc += W[2 * q];
  }
  const double t = block_sum256(acc);
  if (threadIdx.x == 0) parts[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_adj_q0_sum(const double* __restrict__ parts, int n, double* __restrict__ w0) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += parts[i];
  const double t = block_sum256(acc);
  if (threadIdx.x == 0) *w0 = t;
}

// ----- plain FP64, any grid ----------------------------------------------------------------------

template <bool CQ>
__global__ __launch_bounds__(256) void k_adjoint64(QView g, const double* __restrict__ wt, const double* __restrict__ rec,
                                                   int64_t ntri, int64_t qchunk, double* __restrict__ part) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tc = t < ntri ? t : ntri - 1;
  const int chunk = blockIdx.y;
  const int64_t q0 = (int64_t)chunk * qchunk, q1 = min(g.nq, q0 + qchunk);
  const double* r = rec + 8 * tc;
  const double ax = r[0], ay = r[1], az = r[2], cx = r[3], cy = r[4], cz = r[5];
  double gA[3] = {0, 0, 0}, gc[3] = {0, 0, 0};
  for (int64_t q = q0; q < q1; ++q) {
    // q-point data are the same for every lane: scalar loads
    const int64_t p = q % g.npix;
    const double qx = g.qx[p], qy = g.qy[p], zr = g.qz[2 * q], zi = CQ ? g.qz[2 * q + 1] : 0.0;
    const cd w = cd{wt[2 * q], wt[2 * q + 1]};
    // e = exp(i q.c) with complex qz
    double sn, cs;
    sincos(qx * cx + qy * cy + zr * cz, &sn, &cs);
    double m = CQ ? exp(-zi * cz) : 1.0;
    const cd wa = w * cd{m * cs, m * sn};
    const cd qa = cd{qx * ax + qy * ay + zr * az, zi * az};
    const cd wc = mul_i(wa * qa);
    gA[0] += wa.r * qx;
    gA[1] += wa.r * qy;
    gA[2] += wa.r * zr - wa.i * zi;
    gc[0] += wc.r * qx;
    gc[1] += wc.r * qy;
    gc[2] += wc.r * zr - wc.i * zi;
  }
  if (t >= ntri) return;
  double* o = part + 6 * ((int64_t)chunk * ntri + t);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o[k] = gA[k];
    o[3 + k] = gc[k];
  }
}

// Rotates (dL/dA', dL/dc') back by R, adds the q = 0 term and scatters to the caller's order.
__device__ inline void adj_store(const Rot& R, const double* __restrict__ rec, int64_t t, double w0, double gA[3],
                                 double gc[3], const int64_t* __restrict__ src, double* __restrict__ grad) {
  const double* r = rec + 8 * t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    gA[k] += w0 * r[3 + k] / 3.0;
    gc[k] += w0 * r[k] / 3.0;
  }
  double* o = grad + 6 * src[t];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    o[i] = R.m[3 * i] * gA[0] + R.m[3 * i + 1] * gA[1] + R.m[3 * i + 2] * gA[2];
    o[3 + i] = R.m[3 * i] * gc[0] + R.m[3 * i + 1] * gc[1] + R.m[3 * i + 2] * gc[2];
  }
}

__global__ __launch_bounds__(256) void k_adjoint64_finish(const double* __restrict__ part, int nchunks, int64_t ntri,
                                                          const double* __restrict__ rec, Rot R,
                                                          const double* __restrict__ w0p,
                                                          const int64_t* __restrict__ src, double* __restrict__ grad) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  double gA[3] = {0, 0, 0}, gc[3] = {0, 0, 0};
  for (int c = 0; c < nchunks; ++c) {
    const double* o = part + 6 * ((int64_t)c * ntri + t);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      gA[k] += o[k];
      gc[k] += o[3 + k];
    }
  }
  adj_store(R, rec, t, *w0p, gA, gc, src, grad);
}

__global__ __launch_bounds__(256) void k_adj_zero(double* __restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ----- FP32C, any grid (VALU) ----------------------------------------------------------------------
// One workgroup = one 128-triangle block and one chunk of q-points; thread i takes triangle i of the
// block, whose FP32 record holds A' and c' - o_b. Per stage of 128 q-points each thread prepares one
// q-point in LDS: its FP32 components, in turns and as they are, and its weight folded with the block's
// phase, omega(q) = Wt(q) exp(i q . o_b) (the FP64 phase reduced to turns before the hardware sin/cos,
// times exp(-Im(q_z) o_bz)). The inner loop is then FP32 only: exp(i q . (c' - o_b)) from the
// block-relative phase in turns, and every thread reads the same q-point (an LDS broadcast). Sums run
// over one stage in FP32 and are added to FP64 accumulators; the chunk partials are reduced in order by
// k_adjoint64_finish.
constexpr int kAdjQStage = kBlockTri;

template <bool CQ>
__global__ __launch_bounds__(kBlockTri) void k_adjoint32(QView g, const float* __restrict__ wt,
                                                         const float* __restrict__ rec,
                                                         const double* __restrict__ borig, int64_t ntri,
                                                         int64_t qchunk, double* __restrict__ part) {
  __shared__ float4 sq[2 * kAdjQStage];
  __shared__ float2 sqi[CQ ? kAdjQStage : 1];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int64_t t = (int64_t)b * kBlockTri + tid;  // records are padded to whole blocks
  const int chunk = blockIdx.y;
  const int64_t q0 = (int64_t)chunk * qchunk, q1 = min(g.nq, q0 + qchunk);
  const float* r = rec + 8 * t;
  const float ax = r[0], ay = r[1], az = r[2], cx = r[3], cy = r[4], cz = r[5];
  const double ox = borig[4 * b], oy = borig[4 * b + 1], oz = borig[4 * b + 2];
  double gA[3] = {0, 0, 0}, gc[3] = {0, 0, 0};
  for (int64_t s0 = q0; s0 < q1; s0 += kAdjQStage) {
    // this thread's q-point of the stage; past the chunk's end a zero weight and q = 0 add nothing
    const int64_t q = s0 + tid;
    float4 u = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v = u;
    float2 vi = make_float2(0.0f, 0.0f);
    if (q < q1) {
      const int64_t p = q % g.npix;
      const double qx = g.qx[p], qy = g.qy[p], zr = g.qz[2 * q], zi = CQ ? g.qz[2 * q + 1] : 0.0;
      const double tb = (qx * ox + qy * oy + zr * oz) * kInvTwoPi;
      const float fr = (float)(tb - rint(tb));
      double cb = (double)__builtin_amdgcn_cosf(fr), sb = (double)__builtin_amdgcn_sinf(fr);
      if (CQ) {
        const double m = exp(-zi * oz);
        cb *= m;
        sb *= m;
      }
      const double wr = wt[2 * q], wi = wt[2 * q + 1];
      u = make_float4((float)(qx * kInvTwoPi), (float)(qy * kInvTwoPi), (float)(zr * kInvTwoPi),
                      (float)(wr * cb - wi * sb));
      v = make_float4((float)qx, (float)qy, (float)zr, (float)(wr * sb + wi * cb));
      if (CQ) vi = make_float2((float)(zi * 1.4426950408889634), (float)zi);
    }
    __syncthreads();  // every thread has read the previous stage
    sq[2 * tid] = u;
    sq[2 * tid + 1] = v;
    if (CQ) sqi[tid] = vi;
    __syncthreads();
    float fA0 = 0.0f, fA1 = 0.0f, fA2 = 0.0f, fc0 = 0.0f, fc1 = 0.0f, fc2 = 0.0f;
#pragma unroll 4
    for (int k = 0; k < kAdjQStage; ++k) {
      const float4 a = sq[2 * k], e = sq[2 * k + 1];
      const float ph = a.x * cx + a.y * cy + a.z * cz;
      const float sn = __builtin_amdgcn_sinf(ph), cs = __builtin_amdgcn_cosf(ph);
      // wa = omega exp(i q . (c' - o_b)); qa = q . A'; wc = i wa qa
      float war = a.w * cs - e.w * sn, wai = a.w * sn + e.w * cs;
      const float qar = e.x * ax + e.y * ay + e.z * az;
      if (!CQ) {
        const float wcr = -wai * qar;
        fA0 += war * e.x;
        fA1 += war * e.y;
        fA2 += war * e.z;
        fc0 += wcr * e.x;
        fc1 += wcr * e.y;
        fc2 += wcr * e.z;
      } else {
        const float2 ei = sqi[k];
        const float att = __builtin_amdgcn_exp2f(-(ei.x * cz));
        war *= att;
        wai *= att;
        const float qai = ei.y * az;
        const float wcr = -(war * qai + wai * qar), wci = war * qar - wai * qai;
        fA0 += war * e.x;
        fA1 += war * e.y;
        fA2 += war * e.z - wai * ei.y;
        fc0 += wcr * e.x;
        fc1 += wcr * e.y;
        fc2 += wcr * e.z - wci * ei.y;
      }
    }
    gA[0] += fA0;
    gA[1] += fA1;
    gA[2] += fA2;
    gc[0] += fc0;
    gc[1] += fc1;
    gc[2] += fc2;
  }
  if (t >= ntri) return;
  double* o = part + 6 * ((int64_t)chunk * ntri + t);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o[k] = gA[k];
    o[3 + k] = gc[k];
  }
}

// q-chunks of the FP32C any-grid kernel: enough workgroups to fill the chip (about 8192), whole stages
// per chunk, at most one chunk per stage
static void adjoint32_chunks(int64_t nq, int64_t nblocks, int64_t* qchunk, int* nqc) {
  const int64_t stages = ceil_div(nq, kAdjQStage);
  const int64_t want = std::max<int64_t>(1, std::min<int64_t>(stages, ceil_div(8192, std::max<int64_t>(nblocks, 1))));
  *qchunk = ceil_div(stages, want) * kAdjQStage;
  *nqc = (int)ceil_div(nq, *qchunk);
}

// ----- FP32C on product grids (matrix cores) -------------------------------------------------------

__device__ inline f32x16a mfma32a(float a, float b, f32x16a c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

__device__ inline int adj_acc_row(int rr, int h) { return 8 * (rr >> 2) + 4 * h + (rr & 3); }

// One workgroup = (item, f-pass, 128-triangle block, 32 y-columns) over all z; wave w takes the block's
// triangles 32 w .. 32 w + 31. Per chunk of 64 z-rows the workgroup stages the weight tile
// Wt[64 z][32 y] and, per z, the block's reduced base phase q_z o_bz / 2pi and the FP32 axis values in
// LDS, so each weight is read from memory once per block instead of once per wave, and the inner loop
// has no FP64 work: phase = base(z) + qz(z) (c_z - o_bz) / 2pi in turns for the hardware sin/cos.
// MFMA: lane row r of the A operand is triangle t0 + r at z = 2s + h; lane column r of the B operand is
// Wt(z = 2s + h, y0 + r).
constexpr int kAdjZChunk = 64;

struct AdjZ {
  float base, uz, qzr, qzi, attb, az, pad0, pad1;
};

template <bool CQ>
__device__ inline void adj_fetch(const AdjItem& it, int z0, int y0, double oz, int tid, float2 (&w)[8], AdjZ& d) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int idx = tid + 256 * e, zl = idx / kAdjTile, yl = idx % kAdjTile;
    const int z = z0 + zl, y = y0 + yl;
    w[e] = make_float2(0.0f, 0.0f);
    if (z < it.nz && y < it.ncol) w[e] = *reinterpret_cast<const float2*>(it.wt + 2 * ((int64_t)z * it.ncol + y));
  }
  d = AdjZ{};
  const int z = z0 + tid;
  if (tid < kAdjZChunk && z < it.nz) {
    const double qr = it.qz[2 * (int64_t)z * it.ncol], qi = CQ ? it.qz[2 * (int64_t)z * it.ncol + 1] : 0.0;
    const double tb = qr * oz * kInvTwoPi;
    d.base = (float)(tb - rint(tb));
    d.uz = (float)(qr * kInvTwoPi);
    d.qzr = (float)qr;
    d.qzi = (float)qi;
    d.attb = (float)(-qi * oz * 1.4426950408889634);
    d.az = (float)(qi * 1.4426950408889634);
  }
}

template <bool CQ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_adjoint_product32(
    const AdjItem* __restrict__ items, AdjItem single, int upi) {
  // two LDS buffers: chunk k is read from one while chunk k + 1 is written into the other
  __shared__ float2 s_w[2][kAdjZChunk][kAdjTile];
  __shared__ AdjZ s_z[2][kAdjZChunk];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int tid = threadIdx.x;
  const int item = blockIdx.x / upi, ul = blockIdx.x % upi;
  const AdjItem it = items ? items[item] : single;  // workgroup-uniform copy (scalar loads)
  const int yt = ul % it.nyt;
  const int blk = (ul / it.nyt) % it.nblocks;
  const int fp = ul / (it.nyt * it.nblocks);  // 0: f = 1, 1: f = qz, 2: f = qz^2
  const int r = lane & 31, h = lane >> 5;
  const int ncol = it.ncol, nz = it.nz;
  const int64_t t0 = (int64_t)blk * kBlockTri + wave * kAdjTile;
  const float dcz = it.rec32[8 * (t0 + r) + 5];
  const double oz = it.borig[4 * blk + 2];
  const int y0 = yt * kAdjTile;

  float2 wr[8];
  AdjZ dr;
  adj_fetch<CQ>(it, 0, y0, oz, tid, wr, dr);
#pragma unroll
  for (int e = 0; e < 8; ++e) s_w[0][(tid + 256 * e) / kAdjTile][(tid + 256 * e) % kAdjTile] = wr[e];
  if (tid < kAdjZChunk) s_z[0][tid] = dr;
  __syncthreads();

  f32x16a P1, P2, P3, C1, C2, C3;
#pragma unroll
  for (int k = 0; k < 16; ++k) C1[k] = C2[k] = C3[k] = 0.0f;
  const int nchunk = (nz + kAdjZChunk - 1) / kAdjZChunk;
  for (int ck = 0; ck < nchunk; ++ck) {
    const int cur = ck & 1;
    const bool more = ck + 1 < nchunk;
    if (more) adj_fetch<CQ>(it, (ck + 1) * kAdjZChunk, y0, oz, tid, wr, dr);  // in flight during the MFMAs
#pragma unroll
    for (int k = 0; k < 16; ++k) P1[k] = P2[k] = P3[k] = 0.0f;
#pragma unroll 2
    for (int s = 0; s < kAdjZChunk / 2; ++s) {
      const int zl = 2 * s + h;
      const AdjZ d = s_z[cur][zl];
      const float2 w = s_w[cur][zl][r];
      const float ph = __builtin_fmaf(d.uz, dcz, d.base);
      float er = __builtin_amdgcn_cosf(ph), ei = __builtin_amdgcn_sinf(ph);
      if (CQ) {
        const float att = __builtin_amdgcn_exp2f(__builtin_fmaf(-d.az, dcz, d.attb));
        er *= att;
        ei *= att;
      }
      float ar = er, ai = ei;
      if (fp > 0) {
        float xr = d.qzr * er - d.qzi * ei, xi = d.qzr * ei + d.qzi * er;
        if (fp > 1) {
          const float yr = d.qzr * xr - d.qzi * xi, yi = d.qzr * xi + d.qzi * xr;
          xr = yr;
          xi = yi;
        }
        ar = xr;
        ai = xi;
      }
      P1 = mfma32a(ar, w.x, P1);
      P2 = mfma32a(ai, w.y, P2);
      P3 = mfma32a(ar + ai, w.x + w.y, P3);
    }
    // each chunk's FP32 sums (64 z-rows) are added to the unit's FP32 sums
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      C1[k] += P1[k];
      C2[k] += P2[k];
      C3[k] += P3[k];
    }
    if (more) {
      // buffer cur ^ 1 was last read in chunk ck - 1, which every wave finished before the
      // barrier that ended it
#pragma unroll
      for (int e = 0; e < 8; ++e) s_w[cur ^ 1][(tid + 256 * e) / kAdjTile][(tid + 256 * e) % kAdjTile] = wr[e];
      if (tid < kAdjZChunk) s_z[cur ^ 1][tid] = dr;
    }
    __syncthreads();
  }
  // X[t][y] for this lane's column y and 16 triangle rows; contract with g(y) Ey_t(y) over the 32
  // columns of the wave's half
  const int yv = y0 + r;
  const bool yok = yv < ncol;
  const int yc = min(yv, ncol - 1);
  const double qx0 = it.qx[0], qyv = it.qy[yc];
  const int nform = 3 - fp;
  const int form0 = fp == 0 ? 0 : (fp == 1 ? 3 : 5);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int64_t trow = t0 + adj_acc_row(k, h);
    const int64_t tc = min(trow, it.ntri - 1);
    const double ph = (qx0 * it.rec[8 * tc + 3] + qyv * it.rec[8 * tc + 4]) * kInvTwoPi;
    const float fph = (float)(ph - rint(ph));
    const double eyr = (double)__builtin_amdgcn_cosf(fph), eyi = (double)__builtin_amdgcn_sinf(fph);
    const double xr = (double)C1[k] - (double)C2[k];
    const double xi = (double)C3[k] - (double)C1[k] - (double)C2[k];
    double yr = yok ? xr * eyr - xi * eyi : 0.0;
    double yi = yok ? xr * eyi + xi * eyr : 0.0;
    for (int f = 0; f < nform; ++f) {
      double sr = yr, si = yi;
#pragma unroll
      for (int m = 16; m >= 1; m >>= 1) {
        sr += __shfl_xor(sr, m, 32);
        si += __shfl_xor(si, m, 32);
      }
      if (r == 0 && trow < it.ntri) {
        double* o = it.part + 2 * (((int64_t)(form0 + f) * it.nyt + yt) * it.ntri + trow);
        o[0] = sr;
        o[1] = si;
      }
      yr *= qyv;  // next form: one more power of qy
      yi *= qyv;
    }
  }
}

// Sums the forms over y-tiles (in order) and assembles (dL/dA', dL/dc') per triangle.
__global__ __launch_bounds__(256) void k_adjoint_product_finish(const double* __restrict__ part, int nyt, int64_t ntri,
                                                                const double* __restrict__ rec,
                                                                const double* __restrict__ qxp, Rot R,
                                                                const double* __restrict__ w0p,
                                                                const int64_t* __restrict__ src,
                                                                double* __restrict__ grad) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  double gA[3], gc[3];
  adj_assemble(part, nyt, ntri, t, rec, qxp[0], gA, gc);
  adj_store(R, rec, t, *w0p, gA, gc, src, grad);
}

// ----- split fp16 on product grids ----------------------------------------------------------------
//
// The same bilinear forms on v_mfma_f32_32x32x16_f16 (K = 16 z values per instruction, lane half h
// holding z = 8 h .. 8 h + 7), every FP32 operand x split into x_h = fp16(x) and x_l = fp16(x - x_h)
// and each real product taken as x_h y_h + x_l y_h + x_h y_l, as in the forward product kernel.
// Complex products in Gauss's form: P1 = sum Ar Wr, P2 = sum Ai Wi, P3 = sum (Ar + Ai)(Wr + Wi),
// X = (P1 - P2, P3 - P1 - P2), i.e. 9 MFMAs per 16 z values and 32 x 32 tile.
// Weights are split once per evaluation (k_adj_wmax, k_adj_wsplit), scaled per column by 2^sw(y) so
// that |Wr| + |Wi| <= 2^14; the A operands f(z) Ez_t(z) are built per lane and scaled by 2^sa, uniform
// over the workgroup, chosen from the largest |f(z)| times the attenuation bound. X is unscaled by
// 2^-(sa + sw(y)) before the contraction with g(y) Ey_t(y).
//
// A workgroup = (item, f-pass, 128-triangle block, 64 columns): wave w takes triangles 32 w .. 32 w +
// 31 and both 32-column tiles, so each A operand feeds two tiles; the weights of one 16-z step are
// staged in LDS for the 4 waves (double-buffered, one barrier per step). Work units run block-fastest,
// so workgroups running together share the same weights in L2.

constexpr int kAdjZMax = 256;   // z rows per partial column maximum
// NY: 32-column tiles per wave (2 at 2 waves per SIMD, or 4 at 1 with the accumulators in AGPRs)
constexpr int kAdjZTab = 256;   // z values per per-z table (one per thread)

__global__ __launch_bounds__(256) void k_adj_wmax(const float* __restrict__ wt, int64_t wt_stride, int nz, int ncol,
                                                  float* __restrict__ pmax, int64_t pmax_stride) {
  const int y = blockIdx.x * 256 + threadIdx.x, zc = blockIdx.y, item = blockIdx.z;
  if (y >= ncol) return;
  const float* w = wt + item * wt_stride;
  float m = 0.0f;
  const int z1 = min(nz, (zc + 1) * kAdjZMax);
  for (int z = zc * kAdjZMax; z < z1; ++z) {
    const float2 v = *reinterpret_cast<const float2*>(w + 2 * ((int64_t)z * ncol + y));
    m = fmaxf(m, fabsf(v.x) + fabsf(v.y));
  }
  pmax[item * pmax_stride + (int64_t)zc * ncol + y] = m;
}

// exponent e with v <= 2^e (v > 0)
__device__ inline int adj_exp2_bound(float v) {
  int e;
  (void)frexpf(v, &e);
  return e;
}

typedef _Float16 h16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
typedef float fx2 __attribute__((ext_vector_type(2)));

__device__ inline h16x2 a16pk(float a, float b) { return __builtin_convertvector((fx2){a, b}, h16x2); }

__device__ inline h16x8 a16cat(h16x2 a, h16x2 b, h16x2 c, h16x2 d) {
  typedef _Float16 h16x4 __attribute__((ext_vector_type(4)));
  const h16x4 ab = __builtin_shufflevector(a, b, 0, 1, 2, 3);
  const h16x4 cd = __builtin_shufflevector(c, d, 0, 1, 2, 3);
  return __builtin_shufflevector(ab, cd, 0, 1, 2, 3, 4, 5, 6, 7);
}

// 8 FP32 values to their high and low fp16 halves (x - x_h is exact in FP32). neg1 holds -1.0 as a
// value the compiler cannot see, so x + (-1) x_h stays one mixed-precision FMA (v_fma_mix_f32)
// instead of a conversion and a subtraction.
__device__ inline void adj_split8(const float (&x)[8], float neg1, h16x8& hi, h16x8& lo) {
  h16x2 hh[4], ll[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    hh[e] = a16pk(x[2 * e], x[2 * e + 1]);
    ll[e] = a16pk(__builtin_fmaf((float)hh[e][0], neg1, x[2 * e]), __builtin_fmaf((float)hh[e][1], neg1, x[2 * e + 1]));
  }
  hi = a16cat(hh[0], hh[1], hh[2], hh[3]);
  lo = a16cat(ll[0], ll[1], ll[2], ll[3]);
}

__global__ __launch_bounds__(256) void k_adj_wsplit(const float* __restrict__ wt, int64_t wt_stride, int nz, int ncol,
                                                    const float* __restrict__ pmax, int64_t pmax_stride,
                                                    uint4* __restrict__ w16, int64_t w16_stride,
                                                    int* __restrict__ wexp, int64_t wexp_stride) {
  const int y = blockIdx.x * 256 + threadIdx.x, zo = blockIdx.y, item = blockIdx.z;
  if (y >= ncol) return;
  const int nzo = (nz + 7) / 8, nzc = (nz + kAdjZMax - 1) / kAdjZMax;
  float m = 0.0f;
  for (int zc = 0; zc < nzc; ++zc) m = fmaxf(m, pmax[item * pmax_stride + (int64_t)zc * ncol + y]);
  const int sw = m > 0.0f ? max(-100, min(100, 14 - adj_exp2_bound(m))) : 0;
  if (zo == 0) wexp[item * wexp_stride + y] = sw;
  const float sc = __builtin_ldexpf(1.0f, sw);
  const float* w = wt + item * wt_stride;
  float f[2][8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int z = 8 * zo + e;
    float wr = 0.0f, wi = 0.0f;
    if (z < nz) {
      const float2 v = *reinterpret_cast<const float2*>(w + 2 * ((int64_t)z * ncol + y));
      wr = v.x * sc;
      wi = v.y * sc;
    }
    f[0][e] = wr;
    f[1][e] = wi;
  }
  uint4* o = w16 + item * w16_stride;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    h16x8 hi, lo;
    adj_split8(f[k], -1.0f, hi, lo);
    o[((int64_t)(2 * k) * nzo + zo) * ncol + y] = __builtin_bit_cast(uint4, hi);
    o[((int64_t)(2 * k + 1) * nzo + zo) * ncol + y] = __builtin_bit_cast(uint4, lo);
  }
}

// -x for 8 packed fp16 values: flip each sign bit
__device__ inline h16x8 adj_neg8(h16x8 x) {
  uint4 u = __builtin_bit_cast(uint4, x);
  u.x ^= 0x80008000u;
  u.y ^= 0x80008000u;
  u.z ^= 0x80008000u;
  u.w ^= 0x80008000u;
  return __builtin_bit_cast(h16x8, u);
}

__device__ inline f32x16a mfma16a(h16x8 a, h16x8 b, f32x16a c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

template <bool CQ, int NY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_adjoint_product16(
    const AdjItem* __restrict__ items, AdjItem single, int upi) {
  constexpr int kAdjNY = NY;
  constexpr int kAdjYG = kAdjTile * NY;
  // weights of one step: [buffer][part * 2 + half][z octet h][column]; per-z data of 256 z
  __shared__ uint4 s_w[2][kAdjW16Planes][2][kAdjYG];
  __shared__ float s_zb[2][kAdjZTab], s_zu[2][kAdjZTab], s_zfr[2][kAdjZTab], s_zfi[2][kAdjZTab];
  __shared__ float s_zab[2][CQ ? kAdjZTab : 1], s_zaz[2][CQ ? kAdjZTab : 1];
  __shared__ float s_red[256];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int tid = threadIdx.x;
  const int item = blockIdx.x / upi, ul = blockIdx.x % upi;
  const AdjItem it = items ? items[item] : single;  // workgroup-uniform copy (scalar loads)
  const int ncol = it.ncol, nz = it.nz, nzo = it.nzo;
  const int nyg = (ncol + kAdjYG - 1) / kAdjYG;
  const int blk = ul % it.nblocks;
  const int yg = (ul / it.nblocks) % nyg;
  const int fp = ul / (it.nblocks * nyg);  // 0: f = 1, 1: f = qz, 2: f = qz^2
  const int r = lane & 31, h = lane >> 5;
  const int64_t t0 = (int64_t)blk * kBlockTri + wave * kAdjTile;
  const float dcz = it.rec32[8 * (t0 + r) + 5];
  const double oz = it.borig[4 * blk + 2];
  const int y0 = yg * kAdjYG;

  // A scale 2^sa: |f(z)| att <= 2^(14 - sa) / sqrt(2) over all z, with
  // att <= exp(|Im qz| (|oz| + max |c_z - oz|)) over the block's triangles
  float czmax = 0.0f;
  if (CQ) {
    s_red[tid] = tid < kBlockTri ? fabsf(it.rec32[8 * ((int64_t)blk * kBlockTri + tid) + 5]) : 0.0f;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) s_red[tid] = fmaxf(s_red[tid], s_red[tid + w]);
      __syncthreads();
    }
    czmax = (float)fabs(oz) + s_red[0];
    __syncthreads();
  }
  float lmax = -1e30f;  // log2 of |f(z)| att
  for (int z = tid; z < nz; z += 256) {
    const double qr = it.qz[2 * (int64_t)z * ncol], qi = CQ ? it.qz[2 * (int64_t)z * ncol + 1] : 0.0;
    const double a = sqrt(qr * qr + qi * qi);
    const float l = fp == 0 ? 0.0f : (a > 0.0 ? (float)(fp * log2(a)) : -1e30f);
    lmax = fmaxf(lmax, l + (CQ ? (float)(fabs(qi) * 1.4426950408889634) * czmax : 0.0f));
  }
  s_red[tid] = lmax;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) s_red[tid] = fmaxf(s_red[tid], s_red[tid + w]);
    __syncthreads();
  }
  const int sa = s_red[0] > -1e29f ? max(-100, min(100, 14 - (int)ceilf(s_red[0] + 0.5f))) : 0;

  const int nks = (nz + 15) / 16;
  float neg1;
  asm volatile("v_mov_b32 %0, -1.0" : "=v"(neg1));
  // Per-z data of 256 z (16 steps) at a time, one entry per thread, double-buffered: base phase in
  // turns (FP64-reduced), qz / 2pi, f(z) 2^sa, and the attenuation terms (complex qz).
  float zb = 0.f, zu = 0.f, zfr = 0.f, zfi = 0.f, zab = 0.f, zaz = 0.f;
  auto ztable = [&](int chunk) {
    const int z = chunk * kAdjZTab + tid;
    zb = zu = zfr = zfi = zab = zaz = 0.0f;
    if (z < nz) {
      const double qr = it.qz[2 * (int64_t)z * ncol], qi = CQ ? it.qz[2 * (int64_t)z * ncol + 1] : 0.0;
      const double tb = qr * oz * kInvTwoPi;
      zb = (float)(tb - rint(tb));
      zu = (float)(qr * kInvTwoPi);
      double fr = 1.0, fi = 0.0;
      if (fp == 1) {
        fr = qr;
        fi = qi;
      } else if (fp == 2) {
        fr = qr * qr - qi * qi;
        fi = 2.0 * qr * qi;
      }
      zfr = (float)__builtin_ldexp(fr, sa);
      zfi = (float)__builtin_ldexp(fi, sa);
      zab = (float)(-qi * oz * 1.4426950408889634);
      zaz = (float)(qi * 1.4426950408889634);
    }
  };
  auto zstage = [&](int buf) {
    s_zb[buf][tid] = zb;
    s_zu[buf][tid] = zu;
    s_zfr[buf][tid] = zfr;
    s_zfi[buf][tid] = zfi;
    if (CQ) {
      s_zab[buf][tid] = zab;
      s_zaz[buf][tid] = zaz;
    }
  };
  // one step's weights: 4 parts-halves x 2 z octets x kAdjYG columns, kWPT uint4 per thread
  constexpr int kWPT = 2 * kAdjW16Planes * kAdjYG / 256;
  uint4 wv[kWPT];
  auto fetch = [&](int ks) {
#pragma unroll
    for (int q = 0; q < kWPT; ++q) {
      const int i = tid + 256 * q, fh = i / (2 * kAdjYG), hh = (i / kAdjYG) & 1, col = i % kAdjYG;
      const int zo = 2 * ks + hh, y = y0 + col;
      wv[q] = make_uint4(0, 0, 0, 0);
      if (zo < nzo && y < ncol) wv[q] = it.w16[((int64_t)fh * nzo + zo) * ncol + y];
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < kWPT; ++q) {
      const int i = tid + 256 * q, fh = i / (2 * kAdjYG), hh = (i / kAdjYG) & 1, col = i % kAdjYG;
      s_w[buf][fh][hh][col] = wv[q];
    }
  };
  ztable(0);
  zstage(0);
  fetch(0);
  stage(0);
  __syncthreads();

  // P[0] = Re sum_z A B, P[1] = Im sum_z A B: four real products per complex product
  f32x16a P[2][kAdjNY];
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < kAdjNY; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) P[k][j][e] = 0.0f;
  const int nzt = (nz + kAdjZTab - 1) / kAdjZTab;
  for (int ks = 0; ks < nks; ++ks) {
    const int cur = ks & 1;
    const bool more = ks + 1 < nks;
    const int chunk = ks / (kAdjZTab / 16), zc = chunk & 1;
    const bool newz = (ks % (kAdjZTab / 16)) == 0 && chunk + 1 < nzt;
    if (more) fetch(ks + 1);  // in flight during this step's work
    if (newz) ztable(chunk + 1);
    // A operands: triangle t0 + r at z = 16 ks + 8 h + e
    float ar[2][8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int zl = (ks % (kAdjZTab / 16)) * 16 + 8 * h + e;
      const float ph = __builtin_fmaf(s_zu[zc][zl], dcz, s_zb[zc][zl]);
      float c = __builtin_amdgcn_cosf(ph), sn = __builtin_amdgcn_sinf(ph);
      if (CQ) {
        const float att = __builtin_amdgcn_exp2f(__builtin_fmaf(-s_zaz[zc][zl], dcz, s_zab[zc][zl]));
        c *= att;
        sn *= att;
      }
      // f(z) = qz^p is real on real branches
      const float fr = s_zfr[zc][zl], fi = CQ ? s_zfi[zc][zl] : 0.0f;
      const float re = CQ ? fr * c - fi * sn : fr * c, im = CQ ? fr * sn + fi * c : fr * sn;
      ar[0][e] = re;
      ar[1][e] = im;
    }
    h16x8 AH[2], AL[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) adj_split8(ar[k], neg1, AH[k], AL[k]);
    // -Ai: the halves of -x are the negated halves of x
    const h16x8 NH = adj_neg8(AH[1]), NL = adj_neg8(AL[1]);
#pragma unroll
    for (int j = 0; j < kAdjNY; ++j) {
      const h16x8 BRH = __builtin_bit_cast(h16x8, s_w[cur][0][h][kAdjTile * j + r]);
      const h16x8 BRL = __builtin_bit_cast(h16x8, s_w[cur][1][h][kAdjTile * j + r]);
      const h16x8 BIH = __builtin_bit_cast(h16x8, s_w[cur][2][h][kAdjTile * j + r]);
      const h16x8 BIL = __builtin_bit_cast(h16x8, s_w[cur][3][h][kAdjTile * j + r]);
      // Re: Ar Br - Ai Bi; Im: Ar Bi + Ai Br; each real product as hi hi + lo hi + hi lo
      P[0][j] = mfma16a(AH[0], BRH, P[0][j]);
      P[1][j] = mfma16a(AH[0], BIH, P[1][j]);
      P[0][j] = mfma16a(AL[0], BRH, P[0][j]);
      P[1][j] = mfma16a(AL[0], BIH, P[1][j]);
      P[0][j] = mfma16a(AH[0], BRL, P[0][j]);
      P[1][j] = mfma16a(AH[0], BIL, P[1][j]);
      P[0][j] = mfma16a(NH, BIH, P[0][j]);
      P[1][j] = mfma16a(AH[1], BRH, P[1][j]);
      P[0][j] = mfma16a(NL, BIH, P[0][j]);
      P[1][j] = mfma16a(AL[1], BRH, P[1][j]);
      P[0][j] = mfma16a(NH, BIL, P[0][j]);
      P[1][j] = mfma16a(AH[1], BRL, P[1][j]);
    }
    // the other buffers' last reads were in the previous step (weights) or chunk (per-z data), which
    // every wave finished before the barrier that ended it
    if (more) stage(cur ^ 1);
    if (newz) zstage(zc ^ 1);
    __syncthreads();
  }
  // X[t][y] for this lane's column and 16 triangle rows of each tile; contract with g(y) Ey_t(y) over
  // the tile's 32 columns
  const double qx0 = it.qx[0];
  const int nform = 3 - fp;
  const int form0 = fp == 0 ? 0 : (fp == 1 ? 3 : 5);
#pragma unroll
  for (int j = 0; j < kAdjNY; ++j) {
    const int ytile = kAdjNY * yg + j;
    if (ytile >= it.nyt) continue;  // wave-uniform
    const int yv = y0 + kAdjTile * j + r;
    const bool yok = yv < ncol;
    const int yc = min(yv, ncol - 1);
    const double qyv = it.qy[yc];
    const double unscale = __builtin_ldexp(1.0, -(sa + it.wexp[yc]));
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int64_t trow = t0 + adj_acc_row(k, h);
      const int64_t tc = min(trow, it.ntri - 1);
      const double ph = (qx0 * it.rec[8 * tc + 3] + qyv * it.rec[8 * tc + 4]) * kInvTwoPi;
      const float fph = (float)(ph - rint(ph));
      const double eyr = (double)__builtin_amdgcn_cosf(fph), eyi = (double)__builtin_amdgcn_sinf(fph);
      const double xr = (double)P[0][j][k] * unscale, xi = (double)P[1][j][k] * unscale;
      double yr = yok ? xr * eyr - xi * eyi : 0.0;
      double yi = yok ? xr * eyi + xi * eyr : 0.0;
      for (int f = 0; f < nform; ++f) {
        double sr = yr, si = yi;
#pragma unroll
        for (int m = 16; m >= 1; m >>= 1) {
          sr += __shfl_xor(sr, m, 32);
          si += __shfl_xor(si, m, 32);
        }
        if (r == 0 && trow < it.ntri) {
          double* o = it.part + 2 * (((int64_t)(form0 + f) * it.nyt + ytile) * it.ntri + trow);
          o[0] = sr;
          o[1] = si;
        }
        yr *= qyv;  // next form: one more power of qy
        yi *= qyv;
      }
    }
  }
}

int adjoint_product_variant(const Context* c) { return c->tune.adjoint_variant; }

int64_t adjoint_product_units(const AdjItem& it, int variant) {
  if (variant > 0) return 3 * (int64_t)it.nblocks * ceil_div(it.ncol, kAdjTile * variant);
  return 3 * (int64_t)it.nblocks * it.nyt;
}

int64_t adjoint_pmax_count(int nz, int ncol) { return (int64_t)ceil_div(nz, kAdjZMax) * ncol; }

void adjoint_split_weights(hipStream_t s, int nitems, const float* wt, int64_t wt_stride, int nz, int ncol,
                           float* pmax, int64_t pmax_stride, uint4* w16, int64_t w16_stride, int* wexp,
                           int64_t wexp_stride) {
  if (nitems <= 0) return;
  dim3 gm(ceil_div(ncol, 256), ceil_div(nz, kAdjZMax), nitems);
  k_adj_wmax<<<gm, 256, 0, s>>>(wt, wt_stride, nz, ncol, pmax, pmax_stride);
  dim3 gs(ceil_div(ncol, 256), ceil_div(nz, 8), nitems);
  k_adj_wsplit<<<gs, 256, 0, s>>>(wt, wt_stride, nz, ncol, pmax, pmax_stride, w16, w16_stride, wexp, wexp_stride);
}

void adjoint_product_launch(hipStream_t s, const AdjItem* d_items, const AdjItem& single, int upi, int nunits,
                            bool cq, int variant) {
  if (nunits <= 0) return;
  if (variant == 4) {
    if (cq) k_adjoint_product16<true, 4><<<nunits, 256, 0, s>>>(d_items, single, upi);
    else k_adjoint_product16<false, 4><<<nunits, 256, 0, s>>>(d_items, single, upi);
    return;
  }
  if (variant == 2) {
    if (cq) k_adjoint_product16<true, 2><<<nunits, 256, 0, s>>>(d_items, single, upi);
    else k_adjoint_product16<false, 2><<<nunits, 256, 0, s>>>(d_items, single, upi);
    return;
  }
  if (cq) k_adjoint_product32<true><<<nunits, 256, 0, s>>>(d_items, single, upi);
  else k_adjoint_product32<false><<<nunits, 256, 0, s>>>(d_items, single, upi);
}

}  // namespace hgx

using namespace hgx;

extern "C" int hgx_ff_numeric_adjoint(hgx_ctx* h, int mesh_id, int precision, int flags, const double rot[9],
                                      const double* W, double* grad) {
  (void)flags;
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "ff_numeric_adjoint: no q-grid");
  if (mesh_id < 0 || mesh_id >= (int)c->meshes.size() || !c->meshes[mesh_id].live)
    return set_error(c, HGX_ERR_BAD_HANDLE, "ff_numeric_adjoint: unknown mesh %d", mesh_id);
  if (!W || !grad) return set_error(c, HGX_ERR_INVALID_ARG, "ff_numeric_adjoint: null weights or output");
  if (precision != HGX_PREC_FP64 && precision != HGX_PREC_FP32C)
    return set_error(c, HGX_ERR_INVALID_ARG, "ff_numeric_adjoint: bad precision %d", precision);
  const Mesh& m = c->meshes[mesh_id];
  // FP32C on a grid that is not a product grid (the q-space and detector grids): the VALU kernel
  const bool any32 = precision == HGX_PREC_FP32C && !c->product;
  if (int rc = set_device(c)) return rc;
  QView g{c->d_qx, c->d_qy, c->d_qz, c->npix, c->nq};
  Rot R = make_rot(rot);
  const bool cq = !c->qz_real;
  // workspace: rotated records | w0 | weights | partials
  const int nyt = (int)ceil_div(c->ncol, kAdjTile);
  const int64_t qchunk = 4096;
  const int nqc = (int)ceil_div(c->nq, qchunk);
  const size_t rec_b = (size_t)m.ntri * 8 * sizeof(double);
  const bool identity = rot == nullptr || (R.m[0] == 1 && R.m[1] == 0 && R.m[2] == 0 && R.m[3] == 0 &&
                                           R.m[4] == 1 && R.m[5] == 0 && R.m[6] == 0 && R.m[7] == 0 && R.m[8] == 1);
  const size_t r32_b = (precision == HGX_PREC_FP32C && !identity)
                           ? (size_t)m.nblocks * (kBlockTri * 8 * sizeof(float) + 4 * sizeof(double)) : 0;
  const size_t w0_b = 256 + kAdjQ0Parts * sizeof(double);  // w0, then the parts of its sum
  const size_t wt_b = (size_t)c->nq * 2 * (precision == HGX_PREC_FP64 ? sizeof(double) : sizeof(float));
  int64_t qchunk32 = 0;
  int nqc32 = 0;
  if (any32) adjoint32_chunks(c->nq, m.nblocks, &qchunk32, &nqc32);
  const size_t part_b = precision == HGX_PREC_FP64 ? (size_t)nqc * m.ntri * 6 * sizeof(double)
                        : any32                    ? (size_t)nqc32 * m.ntri * 6 * sizeof(double)
                                                   : (size_t)kAdjForms * nyt * m.ntri * 2 * sizeof(double);
  // split-fp16 weights (FP32C with the fp16 matrix cores): column maxima, 12 B per q-point, column scales
  const int variant = precision == HGX_PREC_FP32C && !any32 ? adjoint_product_variant(c) : 0;
  const bool f16 = variant > 0;
  const int nz = c->nbranch * c->nrow;
  const int nzo = (int)ceil_div(nz, 8);
  const size_t pm_b = f16 ? (size_t)adjoint_pmax_count(nz, c->ncol) * sizeof(float) : 0;
  const size_t w16_b = f16 ? (size_t)kAdjW16Planes * nzo * c->ncol * sizeof(uint4) : 0;
  const size_t we_b = f16 ? (size_t)c->ncol * sizeof(int) : 0;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  if (int rc = ensure_work(c, al(rec_b) + w0_b + al(wt_b) + al(part_b) + al(r32_b) + al(pm_b) + al(w16_b) + we_b))
    return rc;
  char* w = static_cast<char*>(c->d_work);
  double* rrec = reinterpret_cast<double*>(w);
  double* w0 = reinterpret_cast<double*>(w + al(rec_b));
  void* wt = w + al(rec_b) + w0_b;
  double* part = reinterpret_cast<double*>(w + al(rec_b) + w0_b + al(wt_b));
  char* r32 = w + al(rec_b) + w0_b + al(wt_b) + al(part_b);
  float* pmax = reinterpret_cast<float*>(r32 + al(r32_b));
  uint4* w16 = reinterpret_cast<uint4*>(r32 + al(r32_b) + al(pm_b));
  int* wexp = reinterpret_cast<int*>(r32 + al(r32_b) + al(pm_b) + al(w16_b));
  k_adj_rotate<<<ceil_div(m.ntri, 256), 256, 0, c->stream>>>(m.d_rec64, m.ntri, R, rrec);
  HGX_CHECK_LAUNCH(c);
  {
    const int nparts = (int)std::max<int64_t>(1, std::min<int64_t>(kAdjQ0Parts, ceil_div(c->nq, 256)));
    const int64_t per = ceil_div(c->nq, nparts);
    double* parts = w0 + 32;
    k_adj_q0<<<nparts, 256, 0, c->stream>>>(g, W, per, parts);
    HGX_CHECK_LAUNCH(c);
    k_adj_q0_sum<<<1, 256, 0, c->stream>>>(parts, nparts, w0);
  }
  HGX_CHECK_LAUNCH(c);
  if (m.ntri_in > m.ntri) {
    k_adj_zero<<<ceil_div(m.ntri_in * 6, 256), 256, 0, c->stream>>>(grad, m.ntri_in * 6);
    HGX_CHECK_LAUNCH(c);
  }
  if (precision == HGX_PREC_FP64) {
    k_adj_weights<double><<<ceil_div(c->nq, 256), 256, 0, c->stream>>>(g, W, static_cast<double*>(wt));
    HGX_CHECK_LAUNCH(c);
    dim3 grid(ceil_div(m.ntri, 256), nqc);
    {
      LaunchTimer lt(c, true);
      if (cq) k_adjoint64<true><<<grid, 256, 0, c->stream>>>(g, static_cast<double*>(wt), rrec, m.ntri, qchunk, part);
      else k_adjoint64<false><<<grid, 256, 0, c->stream>>>(g, static_cast<double*>(wt), rrec, m.ntri, qchunk, part);
    }
    HGX_CHECK_LAUNCH(c);
    k_adjoint64_finish<<<ceil_div(m.ntri, 256), 256, 0, c->stream>>>(part, nqc, m.ntri, rrec, R, w0, m.d_src, grad);
    HGX_CHECK_LAUNCH(c);
    return HGX_OK;
  }
  k_adj_weights<float><<<ceil_div(c->nq, 256), 256, 0, c->stream>>>(g, W, static_cast<float*>(wt));
  HGX_CHECK_LAUNCH(c);
  AdjItem it;
  it.qx = c->d_qx;
  it.qy = c->d_qy;
  it.qz = c->d_qz;
  it.wt = static_cast<const float*>(wt);
  it.rec = rrec;
  if (identity) {
    it.rec32 = m.d_rec32;
    it.borig = m.d_borig;
  } else {
    float* rr = reinterpret_cast<float*>(r32);
    double* ro = reinterpret_cast<double*>(r32 + (size_t)m.nblocks * kBlockTri * 8 * sizeof(float));
    if (int rc = product_rotate_mesh(c, m, R, rr, ro, c->stream)) return rc;
    it.rec32 = rr;
    it.borig = ro;
  }
  if (any32) {
    dim3 grid((unsigned)m.nblocks, (unsigned)nqc32);
    {
      LaunchTimer lt(c, true);
      if (cq) k_adjoint32<true><<<grid, kBlockTri, 0, c->stream>>>(g, static_cast<const float*>(wt), it.rec32, it.borig,
                                                                  m.ntri, qchunk32, part);
      else k_adjoint32<false><<<grid, kBlockTri, 0, c->stream>>>(g, static_cast<const float*>(wt), it.rec32, it.borig,
                                                                 m.ntri, qchunk32, part);
    }
    HGX_CHECK_LAUNCH(c);
    k_adjoint64_finish<<<ceil_div(m.ntri, 256), 256, 0, c->stream>>>(part, nqc32, m.ntri, rrec, R, w0, m.d_src, grad);
    HGX_CHECK_LAUNCH(c);
    return HGX_OK;
  }
  it.part = part;
  it.w16 = w16;
  it.wexp = wexp;
  it.ntri = m.ntri;
  it.nblocks = (int)m.nblocks;
  it.ncol = c->ncol;
  it.nz = nz;
  it.nyt = nyt;
  it.nzo = nzo;
  const int64_t units = adjoint_product_units(it, variant);
  if (units > 2000000000) return set_error(c, HGX_ERR_INVALID_ARG, "ff_numeric_adjoint: problem too large");
  if (f16) {
    adjoint_split_weights(c->stream, 1, static_cast<const float*>(wt), 0, nz, c->ncol, pmax, 0, w16, 0, wexp, 0);
    HGX_CHECK_LAUNCH(c);
  }
  {
    LaunchTimer lt(c, true);
    adjoint_product_launch(c->stream, nullptr, it, (int)units, (int)units, cq, variant);
  }
  HGX_CHECK_LAUNCH(c);
  k_adjoint_product_finish<<<ceil_div(m.ntri, 256), 256, 0, c->stream>>>(part, nyt, m.ntri, rrec, c->d_qx, R, w0,
                                                                         m.d_src, grad);
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

extern "C" int hgx_tune_adjoint(hgx_ctx* h, int operands, int col_tiles) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (operands != HGX_PRODUCT_FP32 && operands != HGX_PRODUCT_F16X2) return HGX_ERR_INVALID_ARG;
  if (operands == HGX_PRODUCT_F16X2 && col_tiles != 2 && col_tiles != 4) return HGX_ERR_INVALID_ARG;
  c->tune.adjoint_variant = operands == HGX_PRODUCT_F16X2 ? col_tiles : 0;
  return HGX_OK;
}
