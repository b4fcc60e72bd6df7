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

namespace hgx {

typedef float f32x16 __attribute__((ext_vector_type(16)));

static int g_product_bpu = 16;
static int g_product_mode = kProductF16x2;

int product_blocks_per_unit() { return g_product_bpu; }
int product_mode() { return g_product_mode; }

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

// accumulator register rr of lane (h = lane / 32) holds tile row 8 (rr / 4) + 4 h + rr % 4
__device__ inline int acc_row(int rr, int h) { return 8 * (rr >> 2) + 4 * h + (rr & 3); }

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
    double* dst = it.nchunks == 1 ? it.out + 2 * q : it.part + 2 * ((int64_t)u.chunk * it.nq + q);
    if (GAUSS) {
      dst[0] = (double)C1[k] - (double)C2[k];
      dst[1] = (double)C3[k] - (double)C1[k] - (double)C2[k];
    } else {
      dst[0] = (double)C1[k];
      dst[1] = (double)C2[k];
    }
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
// K layout of the fp16 MFMA: lane (r, h) feeds row r of A and column r of B at k = 8 h .. 8 h + 7
// (halves in register order), so a lane half owns 8 k-slots; a lane half takes two triangles per
// step and the 12 products per triangle and complex part (4 real products x 3 cross terms) fill
// three instructions per part.

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

typedef float f32x2 __attribute__((ext_vector_type(2)));

// two FP32 values to fp16, rounded to nearest (v_cvt_pk_f16_f32)
__device__ inline f16x2 pk16(float a, float b) { return __builtin_convertvector((f32x2){a, b}, f16x2); }

__device__ inline f16x8 cat4(f16x2 a, f16x2 b, f16x2 c, f16x2 d) {
  typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
  f16x4 ab = __builtin_shufflevector(a, b, 0, 1, 2, 3);
  f16x4 cd = __builtin_shufflevector(c, d, 0, 1, 2, 3);
  return __builtin_shufflevector(ab, cd, 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ inline f32x16 mfma16(f16x8 a, f16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// One triangle's split operands for one lane: z side a[0..3] = (Ez, Z2) high, (Ez, Z2) low;
// y side for the real part b[0..3] = (Y1r, -Y1i), (Eyr, -Eyi) high, then low; for the imaginary part
// c[0..3] = (Y1i, Y1r), (Eyi, Eyr) high, then low.
struct Split16 {
  f16x2 a[4], b[4], c[4];
};

struct Lane16 {
  float ux0, uy, uz;     // q components in turns
  float qx0f, qyf, qzrf, qzif, azf;
  float fy;              // 2^sy
  float fz;              // 2^sz (real qz)
};

template <bool CQ>
__device__ inline void split_triangle(const float4 ra, const float4 rc, const Lane16& L, float by, float bz, float bza,
                                      Split16& o) {
  const float snx = ra.x, sny = ra.y, snz = ra.z, dcx = ra.w, dcy = rc.x, dcz = rc.y;
  const float phy = __builtin_fmaf(L.uy, dcy, __builtin_fmaf(L.ux0, dcx, by));
  const float eyr = __builtin_amdgcn_cosf(phy) * L.fy, eyi = __builtin_amdgcn_sinf(phy) * L.fy;
  const float al = __builtin_fmaf(L.qyf, sny, L.qx0f * snx);
  const float y1r = al * eyr, y1i = al * eyi;
  const float phz = __builtin_fmaf(L.uz, dcz, bz);
  float ezr = __builtin_amdgcn_cosf(phz), ezi = __builtin_amdgcn_sinf(phz);
  float z2r, z2i;
  if (CQ) {
    // bza carries log2 of the block's attenuation and the row scale sz
    const float att = __builtin_amdgcn_exp2f(__builtin_fmaf(-L.azf, dcz, bza));
    ezr *= att;
    ezi *= att;
    const float br = L.qzrf * snz, bi = L.qzif * snz;
    z2r = br * ezr - bi * ezi;
    z2i = br * ezi + bi * ezr;
  } else {
    ezr *= L.fz;
    ezi *= L.fz;
    const float br = L.qzrf * snz;
    z2r = br * ezr;
    z2i = br * ezi;
  }
  o.a[0] = pk16(ezr, ezi);
  o.a[1] = pk16(z2r, z2i);
  o.a[2] = pk16(ezr - (float)o.a[0][0], ezi - (float)o.a[0][1]);
  o.a[3] = pk16(z2r - (float)o.a[1][0], z2i - (float)o.a[1][1]);
  o.b[0] = pk16(y1r, -y1i);
  o.b[1] = pk16(eyr, -eyi);
  // the high halves of -y are -(high halves of y): rounding to nearest is odd
  const float y1rl = y1r - (float)o.b[0][0], y1il = y1i + (float)o.b[0][1];
  const float eyrl = eyr - (float)o.b[1][0], eyil = eyi + (float)o.b[1][1];
  o.b[2] = pk16(y1rl, -y1il);
  o.b[3] = pk16(eyrl, -eyil);
  o.c[0] = pk16(y1i, y1r);
  o.c[1] = pk16(eyi, eyr);
  o.c[2] = pk16(y1il, y1rl);
  o.c[3] = pk16(eyil, eyrl);
}

// smallest e >= 0 with v <= 2^e (v = 0 gives 0)
__device__ inline int bound_exp(float v) {
  if (!(v > 0.0f)) return 0;
  int e;
  (void)frexpf(v, &e);  // v = m 2^e, m in [1/2, 1)
  return max(e, 0);
}

template <bool CQ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_ff_product16(
    const ProductItem* __restrict__ items, const ProductUnit* __restrict__ units, int nunits, ProductItem single,
    int nzt, int nyt) {
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
  const int z0 = u.ztile * kProductTile, y0 = u.ytile * kProductTile;
  const int ncol = it.ncol, nz = it.nz;
  const int zr = min(z0 + r, nz - 1), yc = min(y0 + r, ncol - 1);
  const double qx0 = it.qx[0];
  const double qyv = it.qy[yc];
  const double qzr = it.qz[2 * (int64_t)zr * ncol];
  const double qzi = CQ ? it.qz[2 * (int64_t)zr * ncol + 1] : 0.0;
  Lane16 L;
  L.ux0 = (float)(qx0 * kInvTwoPi);
  L.uy = (float)(qyv * kInvTwoPi);
  L.uz = (float)(qzr * kInvTwoPi);
  L.qx0f = (float)qx0;
  L.qyf = (float)qyv;
  L.qzrf = (float)qzr;
  L.qzif = (float)qzi;
  L.azf = (float)(qzi * 1.4426950408889634);
  // operand scales (see above): every scaled entry at most 2^14
  const int sy = 14 - bound_exp((float)sqrt(qx0 * qx0 + qyv * qyv) * it.snmax);
  const int eb = CQ ? min((int)ceilf(fabsf(L.azf) * it.rmax), 40) : 0;
  const int sz = 14 - bound_exp((float)sqrt(qzr * qzr + qzi * qzi) * it.snmax) - eb;
  L.fy = __builtin_ldexpf(1.0f, sy);
  L.fz = __builtin_ldexpf(1.0f, sz);

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
  f32x16 CR, CI;
#pragma unroll
  for (int k = 0; k < 16; ++k) CR[k] = CI[k] = 0.0f;
  for (int64_t b = b0; b < b1; ++b) {
    f32x16 PR, PI;
#pragma unroll
    for (int k = 0; k < 16; ++k) PR[k] = PI[k] = 0.0f;
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
    const double ox = it.borig[4 * b], oy = it.borig[4 * b + 1], oz = it.borig[4 * b + 2];
    const double ty = (qx0 * ox + qyv * oy) * kInvTwoPi;
    const double tz = qzr * oz * kInvTwoPi;
    const float by = (float)(ty - rint(ty)), bz = (float)(tz - rint(tz));
    const float bza = CQ ? (float)(-qzi * oz * 1.4426950408889634) + (float)sz : 0.0f;
#pragma unroll 1
    for (int p = 0; p < kBlockTri / 4; ++p) {
      // this lane half's triangles 4p + 2h and 4p + 2h + 1
      Split16 T, U;
      split_triangle<CQ>(srec[8 * p + 4 * h], srec[8 * p + 4 * h + 1], L, by, bz, bza, T);
      split_triangle<CQ>(srec[8 * p + 4 * h + 2], srec[8 * p + 4 * h + 3], L, by, bz, bza, U);
      const f16x8 At = cat4(T.a[0], T.a[1], T.a[2], T.a[3]);
      const f16x8 Au = cat4(U.a[0], U.a[1], U.a[2], U.a[3]);
      const f16x8 Atu = cat4(T.a[0], T.a[1], U.a[0], U.a[1]);
      // high x high + low x high of each triangle, then high x low of both
      PR = mfma16(At, cat4(T.b[0], T.b[1], T.b[0], T.b[1]), PR);
      PI = mfma16(At, cat4(T.c[0], T.c[1], T.c[0], T.c[1]), PI);
      PR = mfma16(Au, cat4(U.b[0], U.b[1], U.b[0], U.b[1]), PR);
      PI = mfma16(Au, cat4(U.c[0], U.c[1], U.c[0], U.c[1]), PI);
      PR = mfma16(Atu, cat4(T.b[2], T.b[3], U.b[2], U.b[3]), PR);
      PI = mfma16(Atu, cat4(T.c[2], T.c[3], U.c[2], U.c[3]), PI);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      CR[k] += PR[k];
      CI[k] += PI[k];
    }
  }
  const int yj = y0 + r;
  // every lane takes part in the shuffles; stores are masked after
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int row = acc_row(k, h);
    const int szr = __shfl(sz, row, 64);
    const double unscale = __builtin_ldexp(1.0, -(sy + szr));
    const int zi = z0 + row;
    if (yj >= ncol || zi >= nz) continue;
    const int64_t q = (int64_t)zi * ncol + yj;
    double* dst = it.nchunks == 1 ? it.out + 2 * q : it.part + 2 * ((int64_t)u.chunk * it.nq + q);
    dst[0] = (double)CR[k] * unscale;
    dst[1] = (double)CI[k] * unscale;
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

template <bool CQ>
static void launch16_t(hipStream_t s, const ProductItem* items, const ProductUnit* units, int nunits,
                       const ProductItem& single, int nzt, int nyt) {
  unsigned grid = ceil_div(nunits, kProductWaves);
  k_ff_product16<CQ><<<grid, 256, 0, s>>>(items, units, nunits, single, nzt, nyt);
}

static void launch_any(hipStream_t s, const ProductItem* items, const ProductUnit* units, int nunits,
                       const ProductItem& single, int nzt, int nyt, bool cq, int mode) {
  if (nunits <= 0) return;
  if (mode == kProductF16x2) {
    if (cq) launch16_t<true>(s, items, units, nunits, single, nzt, nyt);
    else launch16_t<false>(s, items, units, nunits, single, nzt, nyt);
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
  int nzt = (int)ceil_div(item.nz, kProductTile), nyt = (int)ceil_div(item.ncol, kProductTile);
  int nunits = nzt * nyt * item.nchunks;  // callers check the count fits
  launch_any(s, nullptr, nullptr, nunits, item, nzt, nyt, cq, mode);
}

}  // namespace hgx

using namespace hgx;

extern "C" int hgx_tune_product(int blocks_per_unit, int complex_mults, int operands) {
  if (blocks_per_unit < 1 || (complex_mults != 3 && complex_mults != 4)) return HGX_ERR_INVALID_ARG;
  if (operands != HGX_PRODUCT_FP32 && operands != HGX_PRODUCT_F16X2) return HGX_ERR_INVALID_ARG;
  if (operands == HGX_PRODUCT_F16X2 && complex_mults != 4) return HGX_ERR_INVALID_ARG;
  g_product_bpu = blocks_per_unit;
  g_product_mode = operands == HGX_PRODUCT_F16X2 ? kProductF16x2 : complex_mults == 3 ? kProductF32Gauss : kProductF32x4;
  return HGX_OK;
}
