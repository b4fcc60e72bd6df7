// ewald_mfma.hip -- FP32C approximate-triangle form factor on Ewald-sphere q-grids, on the matrix cores.
//
// QGrid::create (src/model/qgrid.cpp:76-84) and detector geometries put every pixel on the Ewald
// sphere: q(i, j, b) = (k0 (cos af_i cos tth_j - cos ai), k0 cos af_i sin tth_j, qz_{b,i}), so qx and
// qy depend on the row and the column together and the product-grid factorization of
// product_mfma.hip does not apply. Written as
//     q = (a_i, 0, qz_{b,i}) + c_i v_j,   a_i = k0 (cos af_i - cos ai),  c_i = cos af_i,
//                                        v_j = k0 (cos tth_j - 1, sin tth_j, 0),
// the row part is exact on the z (row) side and the column part is a small vector (|v_j| is about
// |q_y|) scaled per row by c_i, which stays within about 1e-3 of 1 at grazing exit angles. The rows
// of all branches are sorted by c and tiled 64 at a time; with cbar the tile's mid value and
// delta_i = c_i - cbar (|delta| <= D, the tile's half-range),
//     s q.n exp(i q.c) = [c_i Ez(i) . s (v.n) Ebar(j) + beta(i) Ez(i) . Ebar(j)] exp(i delta_i P_j),
//     Ez = exp(i (a_i cx + qz cz)),  beta = s (a_i nx + qz nz),  Ebar = exp(i cbar v_j . c),
//     P_j = v_j . (c - o)  (o: the mesh origin),
// up to the factor exp(i delta_i v_j . o) of the output, applied when the sums are stored. The last
// exponential is expanded to order K: exp(i delta P) = sum_k (delta/D)^k (i D P)^k / k!, so order k
// is the same complex matrix product with row operands (delta_i/D)^k [c Ez, beta Ez] and column
// operands (i D P)^k / k! [s (v.n) Ebar, Ebar]. Its truncation error relative to each term is at most
// x^(K+1) / (K+1)! with x = D max|v| max|c - o|; the host picks K (ewald_orders).
//
// Kernel layout as k_ff_product16s (split fp16, two 32-row tiles per wave, four waves on adjacent
// column tiles, z operands built once per workgroup into LDS and shared, software pipelined): orders
// 0 and 1 carry split fp16 operands (three cross products); orders 2 and up, whose terms are below
// x^2 / 2 of the leading one, carry single fp16 operands (one product).
#include "numeric_core.hpp"
#include "split16.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace hgx {

namespace {

constexpr int kEwRows = 64;  // rows of a tile: two 32-row matrix-core tiles

// The column-side constants of one lane's column.
struct EwCol {
  float ux, uy;  // cbar v / 2 pi (turns per nm)
  float gx, gy;  // v
  float Df;      // the tile's D
  float fy;      // 2^sy
};

// One triangle's column operand pairs for the orders 0 .. NO-1, in the imaginary-part form (Y1i, Y1r),
// (Eyi, Eyr) and the real-part form (Y1r, -Y1i), (Eyr, -Eyi), each built by one conversion (the
// negation is a source modifier); low halves for the first NS orders.
template <int NO, int NS>
struct EwTri {
  f16x2 i0[NO], i1[NO], r0[NO], r1[NO];
  f16x2 li0[NS], li1[NS], lr0[NS], lr1[NS];
};

template <int NO, int NS>
__device__ inline void ew_tri(const float4 ra, const float4 rc, const EwCol& k, float by, float pb, EwTri<NO, NS>& o) {
  const float snx = ra.x, sny = ra.y, dcx = ra.w, dcy = rc.x;
  const float phy = __builtin_fmaf(k.uy, dcy, __builtin_fmaf(k.ux, dcx, by));
  float er = __builtin_amdgcn_cosf(phy) * k.fy, ei = __builtin_amdgcn_sinf(phy) * k.fy;
  const float gam = __builtin_fmaf(k.gy, sny, k.gx * snx);                         // s (v . n)
  const float f = k.Df * __builtin_fmaf(k.gy, dcy, __builtin_fmaf(k.gx, dcx, pb));  // D v . (c - o)
#pragma unroll
  for (int n = 0; n < NO; ++n) {
    if (n > 0) {
      // Ey_n = Ey_{n-1} (i f / n)
      const float s = f * (1.0f / (float)n);
      const float nr = -s * ei, ni = s * er;
      er = nr;
      ei = ni;
    }
    const float yr = gam * er, yi = gam * ei;
    o.i0[n] = pk16(yi, yr);
    o.i1[n] = pk16(ei, er);
    o.r0[n] = pk16n(yr, yi);
    o.r1[n] = pk16n(er, ei);
    if (n < NS) {
      const float dyi = yi - (float)o.i0[n][0], dyr = yr - (float)o.i0[n][1];
      const float dei = ei - (float)o.i1[n][0], der = er - (float)o.i1[n][1];
      o.li0[n] = pk16(dyi, dyr);
      o.li1[n] = pk16(dei, der);
      o.lr0[n] = pk16n(dyr, dyi);
      o.lr1[n] = pk16n(der, dei);
    }
  }
}

// One step's column operands (four triangles, two per lane half): B for the real and imaginary
// parts, per order, high halves and (first NS orders) low halves.
template <int NO, int NS>
struct EwStep {
  f16x8 hre[NO], him[NO];
  f16x8 lre[NS], lim[NS];
};

template <int NO, int NS>
__device__ inline EwStep<NO, NS> ew_step(const float4* __restrict__ rb, int ps, int h, const EwCol& k, float by,
                                         float pb) {
  EwTri<NO, NS> T, U;
  ew_tri<NO, NS>(rb[8 * ps + 4 * h], rb[8 * ps + 4 * h + 1], k, by, pb, T);
  ew_tri<NO, NS>(rb[8 * ps + 4 * h + 2], rb[8 * ps + 4 * h + 3], k, by, pb, U);
  EwStep<NO, NS> o;
#pragma unroll
  for (int n = 0; n < NO; ++n) {
    o.him[n] = cat4(T.i0[n], T.i1[n], U.i0[n], U.i1[n]);
    o.hre[n] = cat4(T.r0[n], T.r1[n], U.r0[n], U.r1[n]);
  }
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    o.lim[n] = cat4(T.li0[n], T.li1[n], U.li0[n], U.li1[n]);
    o.lre[n] = cat4(T.lr0[n], T.lr1[n], U.lr0[n], U.lr1[n]);
  }
  return o;
}

// One work unit per workgroup: 64 sorted rows (tile zt) x 128 columns (column group yg) over one chunk
// of triangle blocks; units run chunk-major, column group fastest.
// PIPE: the column operands of step p + 1 are built during step p and held in registers (one wave per
// SIMD fits); otherwise each step builds its own before its products, which keeps the kernel within
// two waves per SIMD, so two workgroups share a compute unit and cover each other's barriers.
// units == nullptr: the single item `single`, units decoded from blockIdx (chunk-major, column group
// fastest); else unit blockIdx.x of a batch's table, its item from `items`.
template <bool CQ, int K, bool PIPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PIPE ? 1 : 2, 8))) void k_ff_ewald16(
    const EwaldItem* __restrict__ items, const EwaldUnit* __restrict__ units, EwaldItem single, int nyg) {
  constexpr int MZ = 2;
  constexpr int NO = K + 1;
  constexpr int NS = K >= 1 ? 2 : 1;
  constexpr int kSteps = kBlockTri / 4;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int uidx = blockIdx.x;
  int yg, zt, chunk;
  if (units) {
    const EwaldUnit u = units[uidx];
    yg = u.yg;
    zt = u.tile;
    chunk = u.chunk;
  } else {
    yg = uidx % nyg;
    zt = (uidx / nyg) % single.ntile;
    chunk = uidx / (nyg * single.ntile);
  }
  const EwaldItem it = units ? items[units[uidx].item] : single;
  const int r = lane & 31, h = lane >> 5;
  const int ncol = it.ncol;
  const int y0 = (yg * kProductWaves + wave) * kProductTile;
  // columns past the grid are clamped (their sums are not stored)
  const int yc = min(y0 + r, ncol - 1);
  const double cbar = it.tile[2 * zt], D = it.tile[2 * zt + 1];
  const double vx = it.col[2 * yc], vy = it.col[2 * yc + 1];
  EwCol kc;
  kc.ux = (float)(cbar * vx * kInvTwoPi);
  kc.uy = (float)(cbar * vy * kInvTwoPi);
  kc.gx = (float)vx;
  kc.gy = (float)vy;
  kc.Df = (float)D;
  // |s (v.n)| <= |v| max|s n| bounds every column entry (the orders above 0 are smaller)
  const int sy = 14 - bound_exp((float)sqrt(vx * vx + vy * vy) * it.snmax);
  kc.fy = __builtin_ldexpf(1.0f, sy);
  // the row this lane builds z entries for: tile slot `lane` (slots past the last row take the
  // tile's first row; they are not stored)
  const int zslot = it.perm[zt * kEwRows + lane];
  const int zr = zslot >= 0 ? zslot : it.perm[zt * kEwRows];
  const int irow = zr % it.nrow;
  const double ar = it.row[2 * irow], cr = it.row[2 * irow + 1];
  const double qzr = it.qz[2 * (int64_t)zr * ncol];
  const double qzi = CQ ? it.qz[2 * (int64_t)zr * ncol + 1] : 0.0;
  const float dhat = D > 0.0 ? (float)((cr - cbar) / D) : 0.0f;
  const float ua = (float)(ar * kInvTwoPi), uz = (float)(qzr * kInvTwoPi);
  const float arf = (float)ar, crf = (float)cr, qzrf = (float)qzr, qzif = (float)qzi;
  const float azf = (float)(qzi * 1.4426950408889634);  // log2(e) Im(qz): exp(-Im(qz) c) = exp2(-azf c)
  const int eb = CQ ? min((int)ceilf(fabsf(azf) * it.rmax), 40) : 0;
  // |beta| <= |(a, qz)| max|s n|; |c| <= 1 + |c - 1| keeps c Ez within the same bound when c <= 1
  const double rq = sqrt(ar * ar + qzr * qzr + qzi * qzi);
  const int sz = 14 - bound_exp((float)fmax(rq * it.snmax, fabs(cr))) - eb;
  const float fz = __builtin_ldexpf(1.0f, sz);

  // block-origin phases in turns reduced in FP64, and the block's part of P
  auto phases = [&](int64_t b, float& by, float& pb, float& bz, float& bza) {
    const double ox = it.borig[4 * b], oy = it.borig[4 * b + 1], oz = it.borig[4 * b + 2];
    const double ty = cbar * (vx * ox + vy * oy) * kInvTwoPi;
    by = (float)(ty - rint(ty));
    pb = (float)(vx * (ox - it.ox) + vy * (oy - it.oy));
    const double tz = (ar * ox + qzr * oz) * kInvTwoPi;
    bz = (float)(tz - rint(tz));
    bza = CQ ? (float)(-qzi * oz * 1.4426950408889634) + (float)sz : 0.0f;
  };

  const int64_t b0 = (int64_t)chunk * it.bchunk;
  const int nb = (int)(min(it.nblocks, b0 + it.bchunk) - b0);
  __shared__ float4 s_rec[2][2 * kBlockTri];
  // z operands: [buffer][order][row tile][lane half][row] of 16 B, as in k_ff_product16s
  __shared__ uint4 s_zh[2][NO][MZ][2][kProductTile];
  __shared__ uint4 s_zl[2][NS][MZ][2][kProductTile];
  const int hw = wave >> 1, sw = wave & 1;
  const int zm = lane >> 5, zrr = lane & 31;
  auto build = [&](const float4* __restrict__ rb, int ps, float bz, float bza, int buf) {
    const float4 ra = rb[2 * (4 * ps + wave)], rc = rb[2 * (4 * ps + wave) + 1];
    const float snx = ra.x, snz = ra.z, dcx = ra.w, dcz = rc.y;
    const float phz = __builtin_fmaf(ua, dcx, __builtin_fmaf(uz, dcz, bz));
    float ezr = __builtin_amdgcn_cosf(phz), ezi = __builtin_amdgcn_sinf(phz);
    float z2r, z2i;
    const float br = __builtin_fmaf(arf, snx, qzrf * snz);
    if (CQ) {
      const float att = __builtin_amdgcn_exp2f(__builtin_fmaf(-azf, dcz, bza));
      ezr *= att;
      ezi *= att;
      const float bi = qzif * snz;
      z2r = br * ezr - bi * ezi;
      z2i = br * ezi + bi * ezr;
    } else {
      ezr *= fz;
      ezi *= fz;
      z2r = br * ezr;
      z2i = br * ezi;
    }
    float v[4] = {crf * ezr, crf * ezi, z2r, z2i};
#pragma unroll
    for (int n = 0; n < NO; ++n) {
      if (n > 0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] *= dhat;
      }
      const f16x2 h0 = pk16(v[0], v[1]), h1 = pk16(v[2], v[3]);
      uint2 vh;
      vh.x = __builtin_bit_cast(unsigned, h0);
      vh.y = __builtin_bit_cast(unsigned, h1);
      reinterpret_cast<uint2*>(&s_zh[buf][n][zm][hw][zrr])[sw] = vh;
      if (n < NS) {
        const f16x2 l0 = lo16(h0, v[0], v[1]);
        const f16x2 l1 = lo16(h1, v[2], v[3]);
        uint2 vl;
        vl.x = __builtin_bit_cast(unsigned, l0);
        vl.y = __builtin_bit_cast(unsigned, l1);
        reinterpret_cast<uint2*>(&s_zl[buf][n][zm][hw][zrr])[sw] = vl;
      }
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
    float by, pb, bz, bza, byn = 0.0f, pbn = 0.0f, bzn = 0.0f, bzan = 0.0f;
    phases(b0, by, pb, bz, bza);
    build(s_rec[0], 0, bz, bza, 0);
    EwStep<NO, NS> Y;
    if constexpr (PIPE) Y = ew_step<NO, NS>(s_rec[0], 0, h, kc, by, pb);
    for (int lb = 0; lb < nb; ++lb) {
      const int cur = lb & 1;
      const bool more = lb + 1 < nb;
      if (more) phases(b0 + lb + 1, byn, pbn, bzn, bzan);
      f32x16 PR[MZ], PI[MZ];
#pragma unroll
      for (int m = 0; m < MZ; ++m)
#pragma unroll
        for (int k = 0; k < 16; ++k) PR[m][k] = PI[m][k] = 0.0f;
      // step p: products of step p, row operands of step p + 1 (the next block's step 0 after the
      // last step; after the unit's last step they are built from valid records and never read), and
      // the column operands of step p + 1 (PIPE) or of step p
      auto step = [&](int p) {
        const int buf = p & 1;
        const bool wrap = p == kSteps - 1;
        const float4* rb = s_rec[wrap && more ? cur ^ 1 : cur];
        const int ps = wrap ? 0 : p + 1;
        build(rb, ps, wrap ? bzn : bz, wrap ? bzan : bza, buf ^ 1);
        EwStep<NO, NS> Yn;
        if constexpr (PIPE) Yn = ew_step<NO, NS>(rb, ps, h, kc, wrap ? byn : by, wrap ? pbn : pb);
        else Y = ew_step<NO, NS>(s_rec[cur], p, h, kc, by, pb);
#pragma unroll
        for (int m = 0; m < MZ; ++m) {
#pragma unroll
          for (int n = 0; n < NO; ++n) {
            const f16x8 AH = __builtin_bit_cast(f16x8, s_zh[buf][n][m][h][r]);
            PR[m] = mfma16(AH, Y.hre[n], PR[m]);
            PI[m] = mfma16(AH, Y.him[n], PI[m]);
            if (n < NS) {
              const f16x8 AL = __builtin_bit_cast(f16x8, s_zl[buf][n][m][h][r]);
              PR[m] = mfma16(AL, Y.hre[n], PR[m]);
              PI[m] = mfma16(AL, Y.him[n], PI[m]);
              PR[m] = mfma16(AH, Y.lre[n], PR[m]);
              PI[m] = mfma16(AH, Y.lim[n], PI[m]);
            }
          }
        }
        // spread the next step's VALU work between the matrix products
        constexpr int kMfma = MZ * (6 * NS + 2 * (NO - NS));
#pragma unroll
        for (int i = 0; i < kMfma; ++i) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x002, 6, 0);
        }
        if constexpr (PIPE) Y = Yn;
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
          CR[m][k] += PR[m][k];
          CI[m][k] += PI[m][k];
        }
      if (more) {
        by = byn;
        pb = pbn;
        bz = bzn;
        bza = bzan;
      }
    }
  }
  // store: remove the scales 2^(sy + sz), apply exp(i delta v . o); every lane takes part in the
  // shuffles, stores are masked after. Tile row 32 m + row was built by lane 32 m + row.
  const int yj = y0 + r;
  const double vo = vx * it.ox + vy * it.oy;
#pragma unroll
  for (int m = 0; m < MZ; ++m)
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int rho = kProductTile * m + acc_row(k, h);
      const int zs = __shfl(zslot, rho, 64);
      const int szr = __shfl(sz, rho, 64);
      const double cs = __shfl(cr, rho, 64);
      if (yj >= ncol || zs < 0) continue;
      const double unscale = __builtin_ldexp(1.0, -(sy + szr));
      double sn, cn;
      sincos((cs - cbar) * vo, &sn, &cn);
      const double re = (double)CR[m][k] * unscale, im = (double)CI[m][k] * unscale;
      const int64_t q = (int64_t)zs * ncol + yj;
      store_raw(it.nchunks, it.out, it.part, it.part32, (int64_t)chunk * it.nq + q, q, re * cn - im * sn,
                re * sn + im * cn);
    }
}

template <bool CQ, bool PIPE>
void launch_orders(hipStream_t s, const EwaldItem* items, const EwaldUnit* units, const EwaldItem& it, int orders,
                   unsigned grid, int nyg) {
  switch (orders) {
    case 0: k_ff_ewald16<CQ, 0, PIPE><<<grid, 256, 0, s>>>(items, units, it, nyg); break;
    case 1: k_ff_ewald16<CQ, 1, PIPE><<<grid, 256, 0, s>>>(items, units, it, nyg); break;
    case 2: k_ff_ewald16<CQ, 2, PIPE><<<grid, 256, 0, s>>>(items, units, it, nyg); break;
    default: k_ff_ewald16<CQ, 3, PIPE><<<grid, 256, 0, s>>>(items, units, it, nyg); break;
  }
}

void launch_any(hipStream_t s, const EwaldItem* items, const EwaldUnit* units, const EwaldItem& it, int orders,
                unsigned grid, int nyg, bool cq, bool pipe) {
  if (cq) {
    if (pipe) launch_orders<true, true>(s, items, units, it, orders, grid, nyg);
    else launch_orders<true, false>(s, items, units, it, orders, grid, nyg);
  } else {
    if (pipe) launch_orders<false, true>(s, items, units, it, orders, grid, nyg);
    else launch_orders<false, false>(s, items, units, it, orders, grid, nyg);
  }
}

}  // namespace

int ewald_launch(Context* c, const EwaldItem& it, bool cq, int orders) {
  if (orders < 0 || orders > 3) return set_error(c, HGX_ERR_INVALID_ARG, "ewald: %d series orders", orders);
  const int nyg = (int)ceil_div(it.ncol, kProductWaves * kProductTile);
  const int64_t units = (int64_t)nyg * it.ntile * it.nchunks;
  if (units <= 0 || units > INT32_MAX) return set_error(c, HGX_ERR_INVALID_ARG, "ewald: %lld work units", (long long)units);
  launch_any(c->stream, nullptr, nullptr, it, orders, (unsigned)units, nyg, cq, c->tune.ewald_pipe != 0);
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

void ewald_launch_units(hipStream_t s, const EwaldItem* d_items, const EwaldUnit* d_units, int nunits, bool cq,
                        int orders, bool pipelined) {
  if (nunits <= 0) return;
  EwaldItem none{};
  launch_any(s, d_items, d_units, none, orders, (unsigned)nunits, 1, cq, pipelined);
}

void ewald_fill_item(const Context* gc, const Mesh& m, const Rot& R, EwaldItem& it) {
  it.qz = gc->d_qz;
  it.row = gc->ew.d_row;
  it.col = gc->ew.d_col;
  it.tile = gc->ew.d_tile;
  it.perm = gc->ew.d_perm;
  it.nblocks = m.nblocks;
  it.nq = gc->nq;
  it.ncol = gc->ncol;
  it.nrow = gc->nrow;
  it.ntile = gc->ew.ntile;
  it.snmax = (float)m.snmax;
  it.rmax = (float)(std::sqrt(m.origin[0] * m.origin[0] + m.origin[1] * m.origin[1] + m.origin[2] * m.origin[2]) +
                    m.radius);
  // the expansion point: the mesh origin in the rotated frame of the records, R^T o
  it.ox = R.m[0] * m.origin[0] + R.m[3] * m.origin[1] + R.m[6] * m.origin[2];
  it.oy = R.m[1] * m.origin[0] + R.m[4] * m.origin[1] + R.m[7] * m.origin[2];
  it.oz = R.m[2] * m.origin[0] + R.m[5] * m.origin[1] + R.m[8] * m.origin[2];
}

void ewald_free(Context* c) {
  if (c->ew.d_mem) (void)hipFree(c->ew.d_mem);
  c->ew = EwaldGrid{};
}

// Fits q = (a_i + c_i vx_j, c_i vy_j) to the grid arrays (qz has been checked to depend on the
// (branch, row) only). Gauge: c = 1 on the middle row i0 and vx = 0 on that row's column j0 of
// smallest |qy| (the specular column of a detector grid, where v is smallest), so a_i = qx(i, j0),
// c_i = qy(i, j1) / qy(i0, j1) on the column j1 of largest |qy|, vx_j = qx(i0, j) - a_{i0},
// vy_j = qy(i0, j). The fit is kept when every pixel agrees to 1e-10 (1 + max|q|).
int ewald_update(Context* c) {
  ewald_free(c);
  const int nrow = c->nrow, ncol = c->ncol;
  const int64_t npix = c->npix;
  if (nrow < 1 || ncol < 1 || !c->d_qx) return HGX_OK;
  std::vector<double> qx(npix), qy(npix);
  HGX_CHECK_HIP(c, hipMemcpyAsync(qx.data(), c->d_qx, npix * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HGX_CHECK_HIP(c, hipMemcpyAsync(qy.data(), c->d_qy, npix * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  const int i0 = nrow / 2;
  const double* qx0 = qx.data() + (int64_t)i0 * ncol;
  const double* qy0 = qy.data() + (int64_t)i0 * ncol;
  int j0 = 0, j1 = 0;
  for (int j = 1; j < ncol; ++j) {
    if (std::fabs(qy0[j]) < std::fabs(qy0[j0])) j0 = j;
    if (std::fabs(qy0[j]) > std::fabs(qy0[j1])) j1 = j;
  }
  if (!(std::fabs(qy0[j1]) > 0.0) || !std::isfinite(qy0[j1])) return HGX_OK;
  std::vector<double> row(2 * (size_t)nrow), col(2 * (size_t)ncol);
  for (int i = 0; i < nrow; ++i) {
    row[2 * i] = qx[(int64_t)i * ncol + j0];
    row[2 * i + 1] = qy[(int64_t)i * ncol + j1] / qy0[j1];
  }
  for (int j = 0; j < ncol; ++j) {
    col[2 * j] = qx0[j] - qx0[j0];
    col[2 * j + 1] = qy0[j];
  }
  double qmax = 0.0, err = 0.0;
  for (int i = 0; i < nrow; ++i)
    for (int j = 0; j < ncol; ++j) {
      const int64_t p = (int64_t)i * ncol + j;
      qmax = std::max(qmax, std::max(std::fabs(qx[p]), std::fabs(qy[p])));
      const double ex = qx[p] - (row[2 * i] + row[2 * i + 1] * col[2 * j]);
      const double ey = qy[p] - row[2 * i + 1] * col[2 * j + 1];
      err = std::max(err, std::max(std::fabs(ex), std::fabs(ey)));  // a NaN leaves err unchanged...
      if (!(std::isfinite(ex) && std::isfinite(ey))) return HGX_OK;   // ...so non-finite values end the fit
    }
  if (!(err <= 1e-10 * (1.0 + qmax))) return HGX_OK;
  // rows of every branch sorted by c, tiled 64 at a time
  const int nz = c->nbranch * nrow;
  std::vector<int> order(nz);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return row[2 * (x % nrow) + 1] < row[2 * (y % nrow) + 1]; });
  const int ntile = (nz + kEwRows - 1) / kEwRows;
  std::vector<int> perm((size_t)ntile * kEwRows, -1);
  std::vector<double> tile(2 * (size_t)ntile);
  double dmax = 0.0, vmax = 0.0;
  for (int t = 0; t < ntile; ++t) {
    double lo = INFINITY, hi = -INFINITY;
    for (int l = 0; l < kEwRows && t * kEwRows + l < nz; ++l) {
      const int z = order[t * kEwRows + l];
      perm[t * kEwRows + l] = z;
      lo = std::min(lo, row[2 * (z % nrow) + 1]);
      hi = std::max(hi, row[2 * (z % nrow) + 1]);
    }
    tile[2 * t] = 0.5 * (lo + hi);
    tile[2 * t + 1] = 0.5 * (hi - lo);
    dmax = std::max(dmax, tile[2 * t + 1]);
  }
  for (int j = 0; j < ncol; ++j) vmax = std::max(vmax, std::hypot(col[2 * j], col[2 * j + 1]));
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t brow = al(row.size() * sizeof(double)), bcol = al(col.size() * sizeof(double)),
               btile = al(tile.size() * sizeof(double)), bperm = al(perm.size() * sizeof(int));
  void* mem = nullptr;
  if (hipMalloc(&mem, brow + bcol + btile + bperm) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(c, HGX_ERR_OOM, "q-grid: Ewald-grid table allocation failed");
  }
  char* p = static_cast<char*>(mem);
  c->ew.d_mem = mem;
  HGX_CHECK_HIP(c, hipMemcpyAsync(p, row.data(), row.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HGX_CHECK_HIP(c, hipMemcpyAsync(p + brow, col.data(), col.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HGX_CHECK_HIP(c, hipMemcpyAsync(p + brow + bcol, tile.data(), tile.size() * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream));
  HGX_CHECK_HIP(c, hipMemcpyAsync(p + brow + bcol + btile, perm.data(), perm.size() * sizeof(int),
                                  hipMemcpyHostToDevice, c->stream));
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));  // the host tables are released on return
  c->ew.d_row = reinterpret_cast<const double*>(p);
  c->ew.d_col = reinterpret_cast<const double*>(p + brow);
  c->ew.d_tile = reinterpret_cast<const double*>(p + brow + bcol);
  c->ew.d_perm = reinterpret_cast<const int*>(p + brow + bcol + btile);
  c->ew.ntile = ntile;
  c->ew.dmax = dmax;
  c->ew.vmax = vmax;
  c->ew.ok = true;
  return HGX_OK;
}

// Series orders for a mesh of the given radius about its origin: the smallest K whose truncation
// x^(K+1) / (K+1)! plus the single-fp16 rounding of the orders from 2 up (2^-11 of their size, at
// most 2^-11 (x^2/2 + x^3/6)) stays within 3e-8 of each term, x = D max|v| radius.
int ewald_orders(const Context* c, double radius) {
  if (!c->ew.ok) return -1;
  const double x = c->ew.dmax * c->ew.vmax * radius;
  if (!(x >= 0.0) || !std::isfinite(x)) return -1;
  if (x == 0.0) return 0;
  const double tol = 3e-8, u = 1.0 / 2048.0;
  const double x2 = x * x, x3 = x2 * x, x4 = x3 * x;
  if (x2 / 2 <= tol) return 1;
  if (x3 / 6 + u * x2 / 2 <= tol) return 2;
  if (x4 / 24 + u * (x2 / 2 + x3 / 6) <= tol) return 3;
  return -1;
}

}  // namespace hgx

using namespace hgx;

extern "C" int hgx_tune_ewald(hgx_ctx* h, int pipelined) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (pipelined != 0 && pipelined != 1) return HGX_ERR_INVALID_ARG;
  c->tune.ewald_pipe = pipelined;
  return HGX_OK;
}
