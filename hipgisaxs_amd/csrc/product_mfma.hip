// product_mfma.hip -- FP32C approximate-triangle form factor on product q-grids, on the matrix cores.
//
// On a product grid q(zi, yj) = (qx0, qy_j, qz_zi) (zi runs over branch-major rows) the raw sum of
// the approximate method (src/ff/cpu/ff_tri_cpu.cpp:200-246)
//     S(zi, yj) = sum_t (q . s_t n_t) exp(i q . c_t)
// splits per triangle into a z factor and a y factor:
//     q . s n = alpha_t(yj) + beta_t(zi),  alpha = s (qx0 nx + qy ny),  beta = s qz nz
//     exp(i q.c) = Ey_t(yj) Ez_t(zi),      Ey = exp(i (qx0 cx + qy cy)),  Ez = exp(i qz cz)
// so S = sum_t Ez_t(zi) [alpha_t Ey_t](yj) + [beta_t Ez_t](zi) Ey_t(yj): a complex matrix product
// of (z-rows x 2 triangles-terms) by (2 triangle-terms x y-columns). Each wave owns a 32 x 32 tile
// of S and runs it on v_mfma_f32_32x32x2_f32: the MFMA's K = 2 is two triangles (lane half h
// supplies triangle 2p + h), the A operand (lane row r) is the z side, the B operand (lane column
// r) the y side. Every lane computes its own table entries (one sincos pair for y, one for z per
// triangle), so there is no LDS traffic and no barrier. Complex products use either three real
// products (P1 = sum Xr Yr, P2 = sum Xi Yi, P3 = sum (Xr + Xi)(Yr + Yi); Re = P1 - P2,
// Im = P3 - P1 - P2) or four.
//
// Precision: within a 128-triangle block the coordinates are FP32 relative to the block's FP64
// origin o_b. A phase is q.o_b / 2pi (FP64, reduced to [-1/2, 1/2] turns once per block and lane)
// plus q.(c - o_b) / 2pi (FP32, small), fed in turns to the hardware v_sin/v_cos, so each table
// entry is accurate to FP32 rounding whatever |q| |c|. The sums accumulate in FP32 in the
// matrix-core accumulators over one chunk of triangle blocks (16 blocks by default) and chunks are
// added in FP64 by the finishing kernel, which also applies -i/|q|^2 and the small-|q| series
// exactly as for the per-pair path.
#include "numeric_core.hpp"
#include "split16.hpp"

namespace hgx {

int product_blocks_per_unit(const Context* c) { return c->tune.product_bpu; }
int product_mode(const Context* c) {
  const Tuning& t = c->tune;
  if (t.product_operands == HGX_PRODUCT_FP32) return t.product_mults == 3 ? kProductF32Gauss : kProductF32x4;
  if (t.product_shared) return t.product_mz == 4 ? (t.product_two ? kProductF16x2S4 : kProductF16x2S4Flat) : kProductF16x2S2;
  if (t.product_mz == 4) return t.product_two ? kProductF16x2Z4 : kProductF16x2Z4Flat;
  return t.product_mz == 2 ? kProductF16x2Z2 : kProductF16x2;
}

__global__ __launch_bounds__(256) void k_rotate_records(const float* __restrict__ rec, const double* __restrict__ borig,
                                                        int64_t nblocks, Rot R, float* __restrict__ rec_out,
                                                        double* __restrict__ borig_out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nblocks * kBlockTri) {
    const float* a = rec + 8 * t;
    float* o = rec_out + 8 * t;
    // (R^T v)_k = sum_j R[j][k] v_j
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      o[k] = (float)(R.m[k] * a[0] + R.m[3 + k] * a[1] + R.m[6 + k] * a[2]);
      o[3 + k] = (float)(R.m[k] * a[3] + R.m[3 + k] * a[4] + R.m[6 + k] * a[5]);
    }
    o[6] = 0.0f;
    o[7] = 0.0f;
  }
  if (t < nblocks) {
    const double* a = borig + 4 * t;
    double* o = borig_out + 4 * t;
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = R.m[k] * a[0] + R.m[3 + k] * a[1] + R.m[6 + k] * a[2];
    o[3] = 0.0;
  }
}

__device__ inline f32x16 mfma32(float a, float b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

template <bool CQ, bool GAUSS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_ff_product32(const ProductItem* __restrict__ items,
                                                      const ProductUnit* __restrict__ units, int nunits,
                                                      ProductItem single, int nzt, int nyt) {
  // the wave index is uniform across the wave; saying so keeps the unit and item loads scalar
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int uidx = (int)blockIdx.x * kProductWaves + wave;  // host keeps unit counts below 2^31
  if (uidx >= nunits) return;  // whole waves leave; the kernel has no barriers
  ProductUnit u;
  if (units) {
    u = units[uidx];
  } else {
    // single item: unit = (chunk, ztile, ytile), y fastest
    u.item = 0;
    u.ytile = uidx % nyt;
    u.ztile = (uidx / nyt) % nzt;
    u.chunk = uidx / (nyt * nzt);
  }
  // wave-uniform copy (scalar loads)
  const ProductItem it = units ? items[u.item] : single;
  const int r = lane & 31, h = lane >> 5;
  const int z0 = u.ztile * kProductTile, y0 = u.ytile * kProductTile;
  const int ncol = it.ncol, nz = it.nz;
  // this lane's operand row (z) and column (y), clamped: an out-of-range row or column of the
  // operands only feeds the same row or column of the product, which is never stored
  const int zr = min(z0 + r, nz - 1), yc = min(y0 + r, ncol - 1);
  const double qx0 = it.qx[0];
  const double qyv = it.qy[yc];
  const double qzr = it.qz[2 * (int64_t)zr * ncol];
  const double qzi = CQ ? it.qz[2 * (int64_t)zr * ncol + 1] : 0.0;
  const float ux0 = (float)(qx0 * kInvTwoPi), uy = (float)(qyv * kInvTwoPi), uz = (float)(qzr * kInvTwoPi);
  const float qx0f = (float)qx0, qyf = (float)qyv, qzrf = (float)qzr, qzif = (float)qzi;
  const float azf = (float)(qzi * 1.4426950408889634);  // log2(e) Im(qz): exp(-Im(qz) c) = exp2(-azf c)

  const int64_t b0 = (int64_t)u.chunk * it.bchunk;
  const int64_t b1 = min(it.nblocks, b0 + it.bchunk);
  // A block's records (128 x 32 B) are staged per wave in LDS; the next block's are loaded into
  // registers while the current block runs, so global-memory latency stays off the matrix pipe.
  // Only this wave touches its LDS slice, and LDS executes one wave's instructions in order, so
  // wave-scope fences (which keep the compiler from reordering) are all the ordering needed.
  __shared__ float4 s_rec[kProductWaves][2 * kBlockTri];
  float4* const srec = s_rec[wave];
  const float4* gsrc = reinterpret_cast<const float4*>(it.rec + 8 * b0 * kBlockTri) + lane;
  float4 pf0 = {0, 0, 0, 0}, pf1 = pf0, pf2 = pf0, pf3 = pf0;
  if (b0 < b1) {
    pf0 = gsrc[0];
    pf1 = gsrc[64];
    pf2 = gsrc[128];
    pf3 = gsrc[192];
  }
  // Each block's sums accumulate in the matrix-core accumulators (FP32) and are then added to the
  // chunk's FP32 sums, so no FP32 sum absorbs more than 128 triangles or 16 blocks directly;
  // chunks are summed in FP64 by the finishing kernel.
  f32x16 C1, C2, C3;
#pragma unroll
  for (int k = 0; k < 16; ++k) C1[k] = C2[k] = C3[k] = 0.0f;
  for (int64_t b = b0; b < b1; ++b) {
    f32x16 P1, P2, P3;
#pragma unroll
    for (int k = 0; k < 16; ++k) P1[k] = P2[k] = P3[k] = 0.0f;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    srec[lane] = pf0;
    srec[lane + 64] = pf1;
    srec[lane + 128] = pf2;
    srec[lane + 192] = pf3;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (b + 1 < b1) {
      gsrc += 2 * kBlockTri;
      pf0 = gsrc[0];
      pf1 = gsrc[64];
      pf2 = gsrc[128];
      pf3 = gsrc[192];
    }
    // phase of the block origin for this lane's column and row, in turns, reduced in FP64:
    // q . c = q . o_b + q . (c - o_b), the first reduced exactly, the second small
    const double ox = it.borig[4 * b], oy = it.borig[4 * b + 1], oz = it.borig[4 * b + 2];
    const double ty = (qx0 * ox + qyv * oy) * kInvTwoPi;
    const double tz = qzr * oz * kInvTwoPi;
    const float by = (float)(ty - rint(ty)), bz = (float)(tz - rint(tz));
    const float bza = CQ ? (float)(-qzi * oz * 1.4426950408889634) : 0.0f;  // log2 of exp(-Im(qz) oz)
#pragma unroll 2
    for (int p = 0; p < kBlockTri / 2; ++p) {
      // this lane's triangle 2p + h: record float4s 2 (2p + h) and 2 (2p + h) + 1
      const float4 ra = srec[4 * p + 2 * h];
      const float4 rc = srec[4 * p + 2 * h + 1];
      const float snx = ra.x, sny = ra.y, snz = ra.z, dcx = ra.w, dcy = rc.x, dcz = rc.y;
      // y side: Ey = exp(i (qx0 cx + qy cy)), Y1 = alpha Ey, Y2 = Ey
      const float phy = __builtin_fmaf(uy, dcy, __builtin_fmaf(ux0, dcx, by));
      const float eyr = __builtin_amdgcn_cosf(phy), eyi = __builtin_amdgcn_sinf(phy);
      const float al = __builtin_fmaf(qyf, sny, qx0f * snx);
      const float y1r = al * eyr, y1i = al * eyi;
      // z side: Ez = exp(i qz cz), Z1 = Ez, Z2 = beta Ez with beta = qz snz
      const float phz = __builtin_fmaf(uz, dcz, bz);
      float ezr = __builtin_amdgcn_cosf(phz), ezi = __builtin_amdgcn_sinf(phz);
      float z2r, z2i;
      if (CQ) {
        const float att = __builtin_amdgcn_exp2f(__builtin_fmaf(-azf, dcz, bza));
        ezr *= att;
        ezi *= att;
        const float br = qzrf * snz, bi = qzif * snz;
        z2r = br * ezr - bi * ezi;
        z2i = br * ezi + bi * ezr;
      } else {
        const float br = qzrf * snz;
        z2r = br * ezr;
        z2i = br * ezi;
      }
      if (GAUSS) {
        P1 = mfma32(ezr, y1r, P1);
        P1 = mfma32(z2r, eyr, P1);
        P2 = mfma32(ezi, y1i, P2);
        P2 = mfma32(z2i, eyi, P2);
        P3 = mfma32(ezr + ezi, y1r + y1i, P3);
        P3 = mfma32(z2r + z2i, eyr + eyi, P3);
      } else {
        // P1 = real part, P2 = imaginary part
        P1 = mfma32(ezr, y1r, P1);
        P1 = mfma32(-ezi, y1i, P1);
        P1 = mfma32(z2r, eyr, P1);
        P1 = mfma32(-z2i, eyi, P1);
        P2 = mfma32(ezr, y1i, P2);
        P2 = mfma32(ezi, y1r, P2);
        P2 = mfma32(z2r, eyi, P2);
        P2 = mfma32(z2i, eyr, P2);
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      C1[k] += P1[k];
      C2[k] += P2[k];
      if (GAUSS) C3[k] += P3[k];
    }
  }
  // raw sums: in place when the item has one chunk, else into the chunk's partial slot
  const int yj = y0 + r;
  if (yj >= ncol) return;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int zi = z0 + acc_row(k, h);
    if (zi >= nz) continue;
    const int64_t q = (int64_t)zi * ncol + yj;
    const int64_t slot = (int64_t)u.chunk * it.nq + q;
    if (GAUSS)
      store_raw(it.nchunks, it.out, it.part, it.part32, slot, q, (double)C1[k] - (double)C2[k],
                (double)C3[k] - (double)C1[k] - (double)C2[k]);
    else
      store_raw(it.nchunks, it.out, it.part, it.part32, slot, q, (double)C1[k], (double)C2[k]);
  }
}

// ---------------------------------------------------------------------------------------------
// Split-fp16 variant. The same product S = Z Y is run on v_mfma_f32_32x32x16_f16, which does eight
// times the FP32 kernel's multiply-adds per instruction in about half its time. Each FP32 table
// entry x is split into x_h = fp16(x) and x_l = fp16(x - x_h) (both rounded to nearest; x - x_h is
// exact in FP32), and a real product x y is taken as x_h y_h + x_l y_h + x_h y_l: the fp16 products
// are exact in the FP32 accumulators and x_l y_l is below FP32 rounding, so each entry carries
// about 2^-24 relative error, as an FP32 entry does, while its low half stays a normal fp16 number.
// The complex products use four real products: with Z = (Ez, Z2) and Y = (Y1, Ey),
//     Re S = Ezr Y1r - Ezi Y1i + Z2r Eyr - Z2i Eyi,   Im S = Ezr Y1i + Ezi Y1r + Z2r Eyi + Z2i Eyr.
// fp16 has 5 exponent bits, so the operands are scaled by powers of two as high as fp16 allows: a
// row's z entries by 2^sz and a column's y entries by 2^sy, with every entry at most 2^14 (fp16
// holds 65504): |Ey| = 1 and |alpha| <= |q_par| max|s n| give sy = 14 - max(0, ceil log2(|q_par|
// max|s n|)); sz likewise, less a bound on log2 exp(-Im(qz) c_z) <= |Im qz| max|c| log2(e). Small
// triangles' entries then sit far above fp16's subnormal range. The scale of an output (row,
// column) is 2^(sz + sy), the same for every triangle, and is taken off when the sums are stored.
//
// Table work is what bounds this kernel (the VALU work that builds the split operands does not
// overlap the matrix pipe), so each wave owns MZ stacked 32 x 32 tiles (32 MZ rows x 32 columns):
// a lane builds its column's y entries once per triangle and uses them for all MZ row tiles.
// K layout of the fp16 MFMA: lane (r, h) feeds row r of A and column r of B at k = 8 h .. 8 h + 7
// (halves in register order), so a lane half owns 8 k-slots, which hold its two triangles of the
// step (T, U) as [T Ez, T Z2, U Ez, U Z2] on the z side and [T Y1, T Ey, U Y1, U Ey] on the y side.
// The three cross terms are three MFMAs on whole operand registers (high x high, low x high,
// high x low), so no operand is copied: per row tile and complex part, 3 MFMAs per 4 triangles.

// One triangle's y-side operand pairs for this lane's column: for the imaginary part (Y1i, Y1r),
// (Eyi, Eyr) and for the real part (Y1r, -Y1i), (Eyr, -Eyi); high halves h*, low halves l*.
struct YSplit {
  f16x2 hre0, hre1, him0, him1, lre0, lre1, lim0, lim1;
};

// One triangle's z-side operand pairs for one row: (Ezr, Ezi), (Z2r, Z2i), high and low halves.
struct ZSplit {
  f16x2 h0, h1, l0, l1;
};

// y side of a triangle: Ey = exp(i (qx0 cx + qy cy)) 2^sy, Y1 = alpha Ey
__device__ inline void split_y(const float4 ra, const float4 rc, float ux0, float uy, float qx0f, float qyf,
                               float fy, float by, YSplit& o) {
  const float snx = ra.x, sny = ra.y, dcx = ra.w, dcy = rc.x;
  const float phy = __builtin_fmaf(uy, dcy, __builtin_fmaf(ux0, dcx, by));
  const float eyr = __builtin_amdgcn_cosf(phy) * fy, eyi = __builtin_amdgcn_sinf(phy) * fy;
  const float al = __builtin_fmaf(qyf, sny, qx0f * snx);
  const float y1r = al * eyr, y1i = al * eyi;
  o.him0 = pk16(y1i, y1r);
  o.him1 = pk16(eyi, eyr);
  o.hre0 = swap_neg(o.him0);
  o.hre1 = swap_neg(o.him1);
  // the low halves: x - x_h is exact in FP32; -x splits into -x_h, -x_l (rounding to nearest is odd)
  o.lim0 = pk16(y1i - (float)o.him0[0], y1r - (float)o.him0[1]);
  o.lim1 = pk16(eyi - (float)o.him1[0], eyr - (float)o.him1[1]);
  o.lre0 = swap_neg(o.lim0);
  o.lre1 = swap_neg(o.lim1);
}

// z side of a triangle for one row: Ez = exp(i qz cz) 2^sz, Z2 = beta Ez. With complex qz the
// attenuation exp(-Im(qz) cz) and the row scale come in through exp2 (bza carries the block's part
// and sz); with real qz through fz = 2^sz.
template <bool CQ>
__device__ inline void split_z(const float4 ra, const float4 rc, float uz, float qzrf, float qzif, float azf, float fz,
                               float bz, float bza, ZSplit& o) {
  const float snz = ra.z, dcz = rc.y;
  const float phz = __builtin_fmaf(uz, dcz, bz);
  float ezr = __builtin_amdgcn_cosf(phz), ezi = __builtin_amdgcn_sinf(phz);
  float z2r, z2i;
  if (CQ) {
    const float att = __builtin_amdgcn_exp2f(__builtin_fmaf(-azf, dcz, bza));
    ezr *= att;
    ezi *= att;
    const float br = qzrf * snz, bi = qzif * snz;
    z2r = br * ezr - bi * ezi;
    z2i = br * ezi + bi * ezr;
  } else {
    ezr *= fz;
    ezi *= fz;
    const float br = qzrf * snz;
    z2r = br * ezr;
    z2i = br * ezi;
  }
  o.h0 = pk16(ezr, ezi);
  o.h1 = pk16(z2r, z2i);
  o.l0 = lo16(o.h0, ezr, ezi);
  o.l1 = lo16(o.h1, z2r, z2i);
}

// MZ: row tiles per wave; TWO: each 128-triangle block's sums go to separate accumulators and are
// added to the chunk sums after the block (two-level FP32 accumulation) instead of accumulating
// over the whole chunk in place.
template <bool CQ, int MZ, bool TWO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((MZ * (TWO ? 2 : 1) > 4) ? 1 : 2, 8)))
void k_ff_product16(const ProductItem* __restrict__ items, const ProductUnit* __restrict__ units, int nunits,
                    ProductItem single, int nzt, int nyt) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int uidx = (int)blockIdx.x * kProductWaves + wave;
  if (uidx >= nunits) return;  // whole waves leave; the kernel has no barriers
  ProductUnit u;
  if (units) {
    u = units[uidx];
  } else {
    u.item = 0;
    u.ytile = uidx % nyt;
    u.ztile = (uidx / nyt) % nzt;
    u.chunk = uidx / (nyt * nzt);
  }
  const ProductItem it = units ? items[u.item] : single;
  const int r = lane & 31, h = lane >> 5;
  const int z0 = u.ztile * (kProductTile * MZ), y0 = u.ytile * kProductTile;
  const int ncol = it.ncol, nz = it.nz;
  // this lane's column and rows, clamped: a clamped row or column only feeds the same row or
  // column of the product, which is never stored
  const int yc = min(y0 + r, ncol - 1);
  const double qx0 = it.qx[0];
  const double qyv = it.qy[yc];
  const float ux0 = (float)(qx0 * kInvTwoPi), uy = (float)(qyv * kInvTwoPi);
  const float qx0f = (float)qx0, qyf = (float)qyv;
  const int sy = 14 - bound_exp((float)sqrt(qx0 * qx0 + qyv * qyv) * it.snmax);
  const float fy = __builtin_ldexpf(1.0f, sy);
  double qzr[MZ], qzi[MZ];
  float uz[MZ], qzrf[MZ], qzif[MZ], azf[MZ], fz[MZ];
  int sz[MZ];
#pragma unroll
  for (int m = 0; m < MZ; ++m) {
    const int zr = min(z0 + kProductTile * m + r, nz - 1);
    qzr[m] = it.qz[2 * (int64_t)zr * ncol];
    qzi[m] = CQ ? it.qz[2 * (int64_t)zr * ncol + 1] : 0.0;
    uz[m] = (float)(qzr[m] * kInvTwoPi);
    qzrf[m] = (float)qzr[m];
    qzif[m] = (float)qzi[m];
    azf[m] = (float)(qzi[m] * 1.4426950408889634);  // log2(e) Im(qz): exp(-Im(qz) c) = exp2(-azf c)
    const int eb = CQ ? min((int)ceilf(fabsf(azf[m]) * it.rmax), 40) : 0;
    sz[m] = 14 - bound_exp((float)sqrt(qzr[m] * qzr[m] + qzi[m] * qzi[m]) * it.snmax) - eb;
    fz[m] = __builtin_ldexpf(1.0f, sz[m]);
  }

  const int64_t b0 = (int64_t)u.chunk * it.bchunk;
  const int64_t b1 = min(it.nblocks, b0 + it.bchunk);
  // block records staged per wave in LDS, next block prefetched into registers (as in the FP32 kernel)
  __shared__ float4 s_rec[kProductWaves][2 * kBlockTri];
  float4* const srec = s_rec[wave];
  const float4* gsrc = reinterpret_cast<const float4*>(it.rec + 8 * b0 * kBlockTri) + lane;
  float4 pf0 = {0, 0, 0, 0}, pf1 = pf0, pf2 = pf0, pf3 = pf0;
  if (b0 < b1) {
    pf0 = gsrc[0];
    pf1 = gsrc[64];
    pf2 = gsrc[128];
    pf3 = gsrc[192];
  }
  f32x16 CR[MZ], CI[MZ];
#pragma unroll
  for (int m = 0; m < MZ; ++m)
#pragma unroll
    for (int k = 0; k < 16; ++k) CR[m][k] = CI[m][k] = 0.0f;
  for (int64_t b = b0; b < b1; ++b) {
    f32x16 PR[MZ], PI[MZ];
#pragma unroll
    for (int m = 0; m < MZ; ++m)
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        PR[m][k] = TWO ? 0.0f : CR[m][k];
        PI[m][k] = TWO ? 0.0f : CI[m][k];
      }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    srec[lane] = pf0;
    srec[lane + 64] = pf1;
    srec[lane + 128] = pf2;
    srec[lane + 192] = pf3;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (b + 1 < b1) {
      gsrc += 2 * kBlockTri;
      pf0 = gsrc[0];
      pf1 = gsrc[64];
      pf2 = gsrc[128];
      pf3 = gsrc[192];
    }
    // block-origin phases in turns, reduced in FP64 (q.c = q.o_b + q.(c - o_b))
    const double ox = it.borig[4 * b], oy = it.borig[4 * b + 1], oz = it.borig[4 * b + 2];
    const double ty = (qx0 * ox + qyv * oy) * kInvTwoPi;
    const float by = (float)(ty - rint(ty));
    float bz[MZ], bza[MZ];
#pragma unroll
    for (int m = 0; m < MZ; ++m) {
      const double tz = qzr[m] * oz * kInvTwoPi;
      bz[m] = (float)(tz - rint(tz));
      bza[m] = CQ ? (float)(-qzi[m] * oz * 1.4426950408889634) + (float)sz[m] : 0.0f;
    }
#pragma unroll 1
    for (int p = 0; p < kBlockTri / 4; ++p) {
      // this lane half's triangles of the step: T = 4p + 2h, U = 4p + 2h + 1
      const float4 ta = srec[8 * p + 4 * h], tc = srec[8 * p + 4 * h + 1];
      const float4 ua = srec[8 * p + 4 * h + 2], uc = srec[8 * p + 4 * h + 3];
      YSplit YT, YU;
      split_y(ta, tc, ux0, uy, qx0f, qyf, fy, by, YT);
      split_y(ua, uc, ux0, uy, qx0f, qyf, fy, by, YU);
      const f16x8 BHre = cat4(YT.hre0, YT.hre1, YU.hre0, YU.hre1);
      const f16x8 BLre = cat4(YT.lre0, YT.lre1, YU.lre0, YU.lre1);
      const f16x8 BHim = cat4(YT.him0, YT.him1, YU.him0, YU.him1);
      const f16x8 BLim = cat4(YT.lim0, YT.lim1, YU.lim0, YU.lim1);
#pragma unroll
      for (int m = 0; m < MZ; ++m) {
        ZSplit ZT, ZU;
        split_z<CQ>(ta, tc, uz[m], qzrf[m], qzif[m], azf[m], fz[m], bz[m], bza[m], ZT);
        split_z<CQ>(ua, uc, uz[m], qzrf[m], qzif[m], azf[m], fz[m], bz[m], bza[m], ZU);
        const f16x8 AH = cat4(ZT.h0, ZT.h1, ZU.h0, ZU.h1);
        const f16x8 AL = cat4(ZT.l0, ZT.l1, ZU.l0, ZU.l1);
        PR[m] = mfma16(AH, BHre, PR[m]);
        PI[m] = mfma16(AH, BHim, PI[m]);
        PR[m] = mfma16(AL, BHre, PR[m]);
        PI[m] = mfma16(AL, BHim, PI[m]);
        PR[m] = mfma16(AH, BLre, PR[m]);
        PI[m] = mfma16(AH, BLim, PI[m]);
      }
    }
#pragma unroll
    for (int m = 0; m < MZ; ++m)
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        CR[m][k] = TWO ? CR[m][k] + PR[m][k] : PR[m][k];
        CI[m][k] = TWO ? CI[m][k] + PI[m][k] : PI[m][k];
      }
  }
  const int yj = y0 + r;
  // every lane takes part in the shuffles; stores are masked after
#pragma unroll
  for (int m = 0; m < MZ; ++m)
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int row = acc_row(k, h);
      const int szr = __shfl(sz[m], row, 64);
      const double unscale = __builtin_ldexp(1.0, -(sy + szr));
      const int zi = z0 + kProductTile * m + row;
      if (yj >= ncol || zi >= nz) continue;
      const int64_t q = (int64_t)zi * ncol + yj;
      store_raw(it.nchunks, it.out, it.part, it.part32, (int64_t)u.chunk * it.nq + q, q, (double)CR[m][k] * unscale,
                (double)CI[m][k] * unscale);
    }
}

// Shared z operands, software-pipelined. A workgroup's four waves own the same 32 MZ rows and four
// adjacent 32-column tiles (one work unit is a 32 MZ x 128 tile). In step s (four triangles) wave w
// builds the z operands of triangle 4s + w for all the unit's rows (lane l: rows l, l + 64, ...) into
// LDS, and every wave reads its row tiles' A operands from there, so a z entry is built once per
// workgroup instead of once per wave; each wave builds its own column's y operands. The operands of
// step s + 1 (y in registers, z into the other LDS buffer) are built in the same barrier interval as
// the matrix products of step s, which do not depend on them, so the VALU work overlaps the matrix
// core. One barrier per step orders a buffer's writes before its reads and its reads before its next
// writes. Steps run on across block boundaries: block records are double-buffered in LDS and stored
// 32 steps before they are first read.

// one step's y operands: B for the real and imaginary parts, high and low halves
struct YOps {
  f16x8 hre, lre, him, lim;
};

__device__ inline YOps y_step(const float4* __restrict__ rb, int ps, int h, float ux0, float uy, float qx0f, float qyf,
                              float fy, float by) {
  const float4 ta = rb[8 * ps + 4 * h], tc = rb[8 * ps + 4 * h + 1];
  const float4 ua = rb[8 * ps + 4 * h + 2], uc = rb[8 * ps + 4 * h + 3];
  YSplit T, U;
  split_y(ta, tc, ux0, uy, qx0f, qyf, fy, by, T);
  split_y(ua, uc, ux0, uy, qx0f, qyf, fy, by, U);
  YOps o;
  o.hre = cat4(T.hre0, T.hre1, U.hre0, U.hre1);
  o.lre = cat4(T.lre0, T.lre1, U.lre0, U.lre1);
  o.him = cat4(T.him0, T.him1, U.him0, U.him1);
  o.lim = cat4(T.lim0, T.lim1, U.lim0, U.lim1);
  return o;
}

template <bool CQ, int MZ, bool TWO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MZ == 4 && TWO ? 1 : 2, 8)))
void k_ff_product16s(const ProductItem* __restrict__ items, const ProductUnit* __restrict__ units, int nunits,
                     ProductItem single, int nzt, int nyg) {
  static_assert(MZ == 2 || MZ == 4, "the workgroup tile's rows are built 64 per pass");
  constexpr int ZROWS = kProductTile * MZ;  // rows of the unit
  constexpr int RPL = ZROWS / 64;           // rows each lane builds z entries for
  constexpr int kSteps = kBlockTri / 4;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int uidx = blockIdx.x;  // one unit per workgroup
  if (uidx >= nunits) return;   // uniform over the workgroup
  ProductUnit u;
  if (units) {
    u = units[uidx];
  } else {
    u.item = 0;
    u.ytile = uidx % nyg;
    u.ztile = (uidx / nyg) % nzt;
    u.chunk = uidx / (nyg * nzt);
  }
  const ProductItem it = units ? items[u.item] : single;
  const int r = lane & 31, h = lane >> 5;
  const int z0 = u.ztile * ZROWS;
  const int y0 = (u.ytile * kProductWaves + wave) * kProductTile;
  const int ncol = it.ncol, nz = it.nz;
  // columns past the grid are clamped (their sums are not stored)
  const int yc = min(y0 + r, ncol - 1);
  const double qx0 = it.qx[0];
  const double qyv = it.qy[yc];
  const float ux0 = (float)(qx0 * kInvTwoPi), uy = (float)(qyv * kInvTwoPi);
  const float qx0f = (float)qx0, qyf = (float)qyv;
  const int sy = 14 - bound_exp((float)sqrt(qx0 * qx0 + qyv * qyv) * it.snmax);
  const float fy = __builtin_ldexpf(1.0f, sy);
  // the rows this lane builds z entries for: z0 + lane + 64 j
  double qzr[RPL], qzi[RPL];
  float uz[RPL], qzrf[RPL], qzif[RPL], azf[RPL], fz[RPL];
  int sz[RPL];
#pragma unroll
  for (int j = 0; j < RPL; ++j) {
    const int zr = min(z0 + lane + 64 * j, nz - 1);
    qzr[j] = it.qz[2 * (int64_t)zr * ncol];
    qzi[j] = CQ ? it.qz[2 * (int64_t)zr * ncol + 1] : 0.0;
    uz[j] = (float)(qzr[j] * kInvTwoPi);
    qzrf[j] = (float)qzr[j];
    qzif[j] = (float)qzi[j];
    azf[j] = (float)(qzi[j] * 1.4426950408889634);
    const int eb = CQ ? min((int)ceilf(fabsf(azf[j]) * it.rmax), 40) : 0;
    sz[j] = 14 - bound_exp((float)sqrt(qzr[j] * qzr[j] + qzi[j] * qzi[j]) * it.snmax) - eb;
    fz[j] = __builtin_ldexpf(1.0f, sz[j]);
  }
  // block-origin phases in turns, reduced in FP64 (q.c = q.o_b + q.(c - o_b))
  auto phases = [&](int64_t b, float& by, float* bz, float* bza) {
    const double ox = it.borig[4 * b], oy = it.borig[4 * b + 1], oz = it.borig[4 * b + 2];
    const double ty = (qx0 * ox + qyv * oy) * kInvTwoPi;
    by = (float)(ty - rint(ty));
#pragma unroll
    for (int j = 0; j < RPL; ++j) {
      const double tz = qzr[j] * oz * kInvTwoPi;
      bz[j] = (float)(tz - rint(tz));
      bza[j] = CQ ? (float)(-qzi[j] * oz * 1.4426950408889634) + (float)sz[j] : 0.0f;
    }
  };

  const int64_t b0 = (int64_t)u.chunk * it.bchunk;
  const int nb = (int)(min(it.nblocks, b0 + it.bchunk) - b0);
  // records of two blocks (256 float4 each, one per thread)
  __shared__ float4 s_rec[2][2 * kBlockTri];
  // z operands: [buffer][row tile][lane half][row] of 16 B (high halves, low halves): the 8 k-slots of
  // one row for one lane half, i.e. triangles 4s + 2h (bytes 0-7) and 4s + 2h + 1 (bytes 8-15)
  __shared__ uint4 s_zh[2][MZ][2][kProductTile];
  __shared__ uint4 s_zl[2][MZ][2][kProductTile];
  // where this wave's triangle of each step goes: lane half hw, bytes 8 sw .. 8 sw + 7
  const int hw = wave >> 1, sw = wave & 1;
  auto build = [&](const float4* __restrict__ rb, int ps, const float* bz, const float* bza, int buf) {
    const float4 ra = rb[2 * (4 * ps + wave)], rc = rb[2 * (4 * ps + wave) + 1];
#pragma unroll
    for (int j = 0; j < RPL; ++j) {
      ZSplit Z;
      split_z<CQ>(ra, rc, uz[j], qzrf[j], qzif[j], azf[j], fz[j], bz[j], bza[j], Z);
      const int rho = lane + 64 * j, m = rho >> 5, rr = rho & 31;
      uint2 vh, vl;
      vh.x = __builtin_bit_cast(unsigned, Z.h0);
      vh.y = __builtin_bit_cast(unsigned, Z.h1);
      vl.x = __builtin_bit_cast(unsigned, Z.l0);
      vl.y = __builtin_bit_cast(unsigned, Z.l1);
      reinterpret_cast<uint2*>(&s_zh[buf][m][hw][rr])[sw] = vh;
      reinterpret_cast<uint2*>(&s_zl[buf][m][hw][rr])[sw] = vl;
    }
  };

  f32x16 CR[MZ], CI[MZ];
#pragma unroll
  for (int m = 0; m < MZ; ++m)
#pragma unroll
    for (int k = 0; k < 16; ++k) CR[m][k] = CI[m][k] = 0.0f;
  if (nb > 0) {
    const float4* gsrc = reinterpret_cast<const float4*>(it.rec + 8 * b0 * kBlockTri) + threadIdx.x;
    s_rec[0][threadIdx.x] = gsrc[0];
    float4 pf = {0, 0, 0, 0};
    if (nb > 1) pf = gsrc[2 * kBlockTri];
    __syncthreads();
    float by, bz[RPL], bza[RPL], byn = 0.0f, bzn[RPL], bzan[RPL];
    phases(b0, by, bz, bza);
#pragma unroll
    for (int j = 0; j < RPL; ++j) bzn[j] = bzan[j] = 0.0f;
    build(s_rec[0], 0, bz, bza, 0);
    YOps Y = y_step(s_rec[0], 0, h, ux0, uy, qx0f, qyf, fy, by);
    for (int lb = 0; lb < nb; ++lb) {
      const int cur = lb & 1;
      const bool more = lb + 1 < nb;
      if (more) phases(b0 + lb + 1, byn, bzn, bzan);
      f32x16 PR[MZ], PI[MZ];
#pragma unroll
      for (int m = 0; m < MZ; ++m)
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          PR[m][k] = TWO ? 0.0f : CR[m][k];
          PI[m][k] = TWO ? 0.0f : CI[m][k];
        }
      // step p: products of step p, operands of step p + 1 (the next block's step 0 after the last
      // step; after the unit's last step they are built from valid records and never read)
      auto step = [&](int p) {
        const int buf = p & 1;
        f16x8 AH[MZ], AL[MZ];
#pragma unroll
        for (int m = 0; m < MZ; ++m) {
          AH[m] = __builtin_bit_cast(f16x8, s_zh[buf][m][h][r]);
          AL[m] = __builtin_bit_cast(f16x8, s_zl[buf][m][h][r]);
        }
        const bool wrap = p == kSteps - 1;
        const float4* rb = s_rec[wrap && more ? cur ^ 1 : cur];
        const int ps = wrap ? 0 : p + 1;
        float bzs[RPL], bzas[RPL];
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
          bzs[j] = wrap ? bzn[j] : bz[j];
          bzas[j] = wrap ? bzan[j] : bza[j];
        }
        build(rb, ps, bzs, bzas, buf ^ 1);
        const YOps Yn = y_step(rb, ps, h, ux0, uy, qx0f, qyf, fy, wrap ? byn : by);
#pragma unroll
        for (int m = 0; m < MZ; ++m) {
          PR[m] = mfma16(AH[m], Y.hre, PR[m]);
          PI[m] = mfma16(AH[m], Y.him, PI[m]);
          PR[m] = mfma16(AL[m], Y.hre, PR[m]);
          PI[m] = mfma16(AL[m], Y.him, PI[m]);
          PR[m] = mfma16(AH[m], Y.lre, PR[m]);
          PI[m] = mfma16(AH[m], Y.lim, PI[m]);
        }
        // spread the next step's VALU work between the matrix products (one MFMA, then a few VALU
        // instructions, and so on), so it issues while the matrix core is busy
#pragma unroll
        for (int i = 0; i < 6 * MZ; ++i) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x002, MZ == 2 ? 7 : 4, 0);
        }
        Y = Yn;
      };
      __syncthreads();
      if (more) {
        // the next block's records: their buffer's last reads were in the previous block
        s_rec[cur ^ 1][threadIdx.x] = pf;
        if (lb + 2 < nb) pf = gsrc[2 * (lb + 2) * kBlockTri];
      }
      step(0);
#pragma unroll 1
      for (int p = 1; p < kSteps; ++p) {
        __syncthreads();
        step(p);
      }
#pragma unroll
      for (int m = 0; m < MZ; ++m)
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          CR[m][k] = TWO ? CR[m][k] + PR[m][k] : PR[m][k];
          CI[m][k] = TWO ? CI[m][k] + PI[m][k] : PI[m][k];
        }
      if (more) {
        by = byn;
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
          bz[j] = bzn[j];
          bza[j] = bzan[j];
        }
      }
    }
  }
  const int yj = y0 + r;
  // every lane takes part in the shuffles; stores are masked after. Row 32 m + row of the unit was
  // built by lane 32 (m & 1) + row, as its row j = m / 2.
#pragma unroll
  for (int m = 0; m < MZ; ++m)
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int row = acc_row(k, h);
      const int szr = __shfl(sz[m / 2], 32 * (m & 1) + row, 64);
      const double unscale = __builtin_ldexp(1.0, -(sy + szr));
      const int zi = z0 + kProductTile * m + row;
      if (yj >= ncol || zi >= nz) continue;
      const int64_t q = (int64_t)zi * ncol + yj;
      store_raw(it.nchunks, it.out, it.part, it.part32, (int64_t)u.chunk * it.nq + q, q, (double)CR[m][k] * unscale,
                (double)CI[m][k] * unscale);
    }
}

int product_rotate_mesh(Context* c, const Mesh& m, const Rot& R, float* rec_out, double* borig_out,
                        hipStream_t s) {
  int64_t n = m.nblocks * kBlockTri;
  k_rotate_records<<<ceil_div(n, 256), 256, 0, s>>>(m.d_rec32, m.d_borig, m.nblocks, R, rec_out, borig_out);
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

template <bool CQ, bool G>
static void launch_t(hipStream_t s, const ProductItem* items, const ProductUnit* units, int nunits,
                     const ProductItem& single, int nzt, int nyt) {
  unsigned grid = ceil_div(nunits, kProductWaves);
  k_ff_product32<CQ, G><<<grid, 256, 0, s>>>(items, units, nunits, single, nzt, nyt);
}

template <bool CQ, int MZ, bool TWO>
static void launch16_t(hipStream_t s, const ProductItem* items, const ProductUnit* units, int nunits,
                       const ProductItem& single, int nzt, int nyt) {
  unsigned grid = ceil_div(nunits, kProductWaves);
  k_ff_product16<CQ, MZ, TWO><<<grid, 256, 0, s>>>(items, units, nunits, single, nzt, nyt);
}

template <bool CQ, int MZ, bool TWO>
static void launch16s_t(hipStream_t s, const ProductItem* items, const ProductUnit* units, int nunits,
                        const ProductItem& single, int nzt, int nyg) {
  k_ff_product16s<CQ, MZ, TWO><<<nunits, 256, 0, s>>>(items, units, nunits, single, nzt, nyg);
}

template <bool CQ>
static void launch16_mode(hipStream_t s, const ProductItem* items, const ProductUnit* units, int nunits,
                          const ProductItem& single, int nzt, int nyt, int mode) {
  switch (mode) {
    case kProductF16x2S2: launch16s_t<CQ, 2, true>(s, items, units, nunits, single, nzt, nyt); break;
    case kProductF16x2S4Flat: launch16s_t<CQ, 4, false>(s, items, units, nunits, single, nzt, nyt); break;
    case kProductF16x2S4: launch16s_t<CQ, 4, true>(s, items, units, nunits, single, nzt, nyt); break;
    case kProductF16x2Z2: launch16_t<CQ, 2, true>(s, items, units, nunits, single, nzt, nyt); break;
    case kProductF16x2Z4: launch16_t<CQ, 4, true>(s, items, units, nunits, single, nzt, nyt); break;
    case kProductF16x2Z4Flat: launch16_t<CQ, 4, false>(s, items, units, nunits, single, nzt, nyt); break;
    default: launch16_t<CQ, 1, true>(s, items, units, nunits, single, nzt, nyt); break;
  }
}

static void launch_any(hipStream_t s, const ProductItem* items, const ProductUnit* units, int nunits,
                       const ProductItem& single, int nzt, int nyt, bool cq, int mode) {
  if (nunits <= 0) return;
  if (product_is_f64(mode)) {
    product64_launch(s, items, units, nunits, single, nzt, nyt, cq, mode == kProductF64Gauss);
  } else if (product_is_f16(mode)) {
    if (cq) launch16_mode<true>(s, items, units, nunits, single, nzt, nyt, mode);
    else launch16_mode<false>(s, items, units, nunits, single, nzt, nyt, mode);
  } else if (cq) {
    if (mode == kProductF32Gauss) launch_t<true, true>(s, items, units, nunits, single, nzt, nyt);
    else launch_t<true, false>(s, items, units, nunits, single, nzt, nyt);
  } else {
    if (mode == kProductF32Gauss) launch_t<false, true>(s, items, units, nunits, single, nzt, nyt);
    else launch_t<false, false>(s, items, units, nunits, single, nzt, nyt);
  }
}

void product_launch_units(hipStream_t s, const ProductItem* d_items, const ProductUnit* d_units, int nunits,
                          bool cq, int mode) {
  ProductItem none{};
  launch_any(s, d_items, d_units, nunits, none, 1, 1, cq, mode);
}

void product_launch_single(hipStream_t s, const ProductItem& item, bool cq, int mode) {
  int nzt = (int)ceil_div(item.nz, product_zrows(mode)), nyt = (int)ceil_div(item.ncol, product_ycols(mode));
  int nunits = nzt * nyt * item.nchunks;  // callers check the count fits
  launch_any(s, nullptr, nullptr, nunits, item, nzt, nyt, cq, mode);
}

}  // namespace hgx

using namespace hgx;

extern "C" int hgx_tune_product(hgx_ctx* h, int blocks_per_unit, int complex_mults, int operands) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (blocks_per_unit < 1 || (complex_mults != 3 && complex_mults != 4)) return HGX_ERR_INVALID_ARG;
  if (operands != HGX_PRODUCT_FP32 && operands != HGX_PRODUCT_F16X2) return HGX_ERR_INVALID_ARG;
  if (operands == HGX_PRODUCT_F16X2 && complex_mults != 4) return HGX_ERR_INVALID_ARG;
  c->tune.product_bpu = blocks_per_unit;
  c->tune.product_operands = operands;
  c->tune.product_mults = complex_mults;
  return HGX_OK;
}

extern "C" int hgx_tune_product64(hgx_ctx* h, int complex_mults) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (complex_mults != 3 && complex_mults != 4) return HGX_ERR_INVALID_ARG;
  c->tune.product64_mults = complex_mults;
  return HGX_OK;
}

extern "C" int hgx_tune_product_tile(hgx_ctx* h, int row_tiles, int two_level, int shared_rows) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!(row_tiles == 1 || row_tiles == 2 || row_tiles == 4)) return HGX_ERR_INVALID_ARG;
  if (!two_level && row_tiles != 4) return HGX_ERR_INVALID_ARG;
  if (shared_rows && row_tiles == 1) return HGX_ERR_INVALID_ARG;
  c->tune.product_mz = row_tiles;
  c->tune.product_two = two_level ? 1 : 0;
  c->tune.product_shared = shared_rows ? 1 : 0;
  return HGX_OK;
}
