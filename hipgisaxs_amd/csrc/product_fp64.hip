// product_fp64.hip -- plain FP64 approximate-triangle form factor on product q-grids, on the FP64
// matrix cores (v_mfma_f64_16x16x4_f64).
//
// The same factorization as the FP32C product kernels (product_mfma.hip): on a product grid
// q(zi, yj) = (qx0, qy_j, qz_zi) the raw sum of the approximate method
// (src/ff/cpu/ff_tri_cpu.cpp:200-246) is
//     S(zi, yj) = sum_t Ez_t(zi) Y1_t(yj) + Z2_t(zi) Ey_t(yj),
//     Ez = exp(i qz cz), Z2 = qz s nz Ez, Ey = exp(i (qx0 cx + qy cy)), Y1 = s (qx0 nx + qy ny) Ey,
// a complex matrix product (z-rows x triangle terms) . (triangle terms x y-columns). Here every table
// entry and every product is FP64, with the triangles' absolute FP64 coordinates: this is the plain
// path on the tiling of the fast one, so the FP32C kernels can be compared with an FP64 kernel that
// does the same algorithmic work.
//
// Tiling. A workgroup (4 waves) owns 64 z-rows x 64 y-columns of one item over one chunk of
// triangles; each wave owns a 32 x 32 quadrant as 2 x 2 tiles of 16 x 16. Per step of 8 triangles
// the workgroup builds the z entries (Ez, Z2) of its 64 rows and the y entries (Y1, Ey) of its 64
// columns into LDS (each thread: one triangle, two rows and two columns: FP64 sincos), double
// buffered with one barrier per step, and every wave reads its MFMA operands from there.
//
// Operands of v_mfma_f64_16x16x4_f64: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16];
// lane l holds D[4 r + l / 16][l % 16] in accumulator element r (measured on gfx950). Complex products in four real
// products (K = 4 = the two terms x real/imaginary parts of one triangle):
//     Re S += [Ezr, -Ezi, Z2r, -Z2i] . [Y1r, Y1i, Eyr, Eyi],  Im S += [Ezr, Ezi, Z2r, Z2i] . [Y1i, Y1r, Eyi, Eyr]
// (two MFMAs per triangle and 16 x 16 tile: 16 flop per (q, triangle) pair), or in Gauss's three
// (K = 4 = two triangles x two terms): P1 = sum Xr Yr, P2 = sum Xi Yi, P3 = sum (Xr + Xi)(Yr + Yi),
// Re = P1 - P2, Im = P3 - P1 - P2 (three MFMAs per two triangles: 12 flop per pair).
#include "numeric_core.hpp"

namespace hgx {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kP64Tile = 64;  // rows and columns of one workgroup's tile
constexpr int kP64Step = 8;   // triangles per LDS step

// one table entry: z side (Ez, Z2), y side (Y1, Ey)
struct alignas(32) P64Entry {
  double ar, ai, br, bi;
};

__device__ inline f64x4 mfma64(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__global__ void k_rotate_records64(const double* __restrict__ rec, int64_t ntri, Rot R, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  const double* a = rec + 8 * t;
  double* o = out + 8 * t;
  // (R^T v)_k = sum_j R[j][k] v_j
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o[k] = R.m[k] * a[0] + R.m[3 + k] * a[1] + R.m[6 + k] * a[2];
    o[3 + k] = R.m[k] * a[3] + R.m[3 + k] * a[4] + R.m[6 + k] * a[5];
  }
  o[6] = 0.0;
  o[7] = 0.0;
}

template <bool CQ, bool GAUSS>
__global__ __launch_bounds__(256) void k_ff_product64(const ProductItem* __restrict__ items,
                                                      const ProductUnit* __restrict__ units, int nunits,
                                                      ProductItem single, int nzt, int nyt) {
  const int uidx = blockIdx.x;
  if (uidx >= nunits) return;  // whole workgroups leave, before any barrier
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
  const ProductItem it = units ? items[u.item] : single;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int ncol = it.ncol, nz = it.nz;
  const int z0 = u.ztile * kP64Tile, y0 = u.ytile * kP64Tile;

  // builder role: triangle bs of each step, rows and columns be, be + 1 of the tile (clamped: an
  // out-of-range row or column only feeds outputs that are never stored)
  const int bs = tid >> 5, be = (tid & 31) * 2;
  const double qx0 = it.qx[0];
  double qzr[2], qzi[2], qyv[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int64_t zr = min(z0 + be + e, nz - 1), yc = min(y0 + be + e, ncol - 1);
    qzr[e] = it.qz[2 * zr * ncol];
    qzi[e] = CQ ? it.qz[2 * zr * ncol + 1] : 0.0;
    qyv[e] = it.qy[yc];
  }

  __shared__ P64Entry s_z[2][kP64Step][kP64Tile];
  __shared__ P64Entry s_y[2][kP64Step][kP64Tile];

  const int64_t t0 = (int64_t)u.chunk * it.bchunk * kBlockTri;
  const int64_t t1 = min(it.ntri, t0 + it.bchunk * kBlockTri);
  const int nsteps = t1 > t0 ? (int)((t1 - t0 + kP64Step - 1) / kP64Step) : 0;

  auto build = [&](int step, int buf) {
    const int64_t t = t0 + (int64_t)step * kP64Step + bs;
    double snx = 0, sny = 0, snz = 0, cx = 0, cy = 0, cz = 0;  // past the chunk: a zero record adds nothing
    if (t < t1) {
      const double* r = it.rec64 + 8 * t;
      snx = r[0];
      sny = r[1];
      snz = r[2];
      cx = r[3];
      cy = r[4];
      cz = r[5];
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      // z side: Ez = exp(i qz cz), Z2 = qz snz Ez
      double s, c;
      sincos(qzr[e] * cz, &s, &c);
      if (CQ) {
        const double m = exp(-qzi[e] * cz);
        s *= m;
        c *= m;
      }
      const double br = qzr[e] * snz, bi = qzi[e] * snz;
      s_z[buf][bs][be + e] = P64Entry{c, s, br * c - bi * s, br * s + bi * c};
      // y side: Ey = exp(i (qx0 cx + qy cy)), Y1 = alpha Ey
      sincos(qx0 * cx + qyv[e] * cy, &s, &c);
      const double al = qx0 * snx + qyv[e] * sny;
      s_y[buf][bs][be + e] = P64Entry{al * c, al * s, c, s};
    }
  };

  // this wave's 32 x 32 quadrant: tiles (mz, my) at rows wz + 16 mz, columns wy + 16 my
  const int wz = (wave >> 1) * 32, wy = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  constexpr int NACC = GAUSS ? 3 : 2;
  f64x4 acc[2][2][NACC];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int n = 0; n < NACC; ++n) acc[a][b][n] = (f64x4){0.0, 0.0, 0.0, 0.0};

  if (nsteps > 0) build(0, 0);
  __syncthreads();
  for (int step = 0; step < nsteps; ++step) {
    const int buf = step & 1;
    if (step + 1 < nsteps) build(step + 1, buf ^ 1);
    if (!GAUSS) {
#pragma unroll 2
      for (int tt = 0; tt < kP64Step; ++tt) {
        double av[2];
#pragma unroll
        for (int mz = 0; mz < 2; ++mz)
          av[mz] = reinterpret_cast<const double*>(&s_z[buf][tt][wz + 16 * mz + li])[lk];
        const double sgn = (lk & 1) ? -1.0 : 1.0;
#pragma unroll
        for (int my = 0; my < 2; ++my) {
          const double* yb = reinterpret_cast<const double*>(&s_y[buf][tt][wy + 16 * my + li]);
          const double bre = yb[lk], bim = yb[lk ^ 1];
#pragma unroll
          for (int mz = 0; mz < 2; ++mz) {
            acc[mz][my][0] = mfma64(sgn * av[mz], bre, acc[mz][my][0]);
            acc[mz][my][1] = mfma64(av[mz], bim, acc[mz][my][1]);
          }
        }
      }
    } else {
      // k slot: triangle 2p + (lk >> 1), term lk & 1 (0: Ez . Y1, 1: Z2 . Ey)
      const int tsel = lk >> 1, term = lk & 1;
#pragma unroll 2
      for (int p = 0; p < kP64Step / 2; ++p) {
        double xr[2], xi[2];
#pragma unroll
        for (int mz = 0; mz < 2; ++mz) {
          const double* zb = reinterpret_cast<const double*>(&s_z[buf][2 * p + tsel][wz + 16 * mz + li]) + 2 * term;
          xr[mz] = zb[0];
          xi[mz] = zb[1];
        }
#pragma unroll
        for (int my = 0; my < 2; ++my) {
          const double* yb = reinterpret_cast<const double*>(&s_y[buf][2 * p + tsel][wy + 16 * my + li]) + 2 * term;
          const double yr = yb[0], yi = yb[1];
#pragma unroll
          for (int mz = 0; mz < 2; ++mz) {
            acc[mz][my][0] = mfma64(xr[mz], yr, acc[mz][my][0]);
            acc[mz][my][1] = mfma64(xi[mz], yi, acc[mz][my][1]);
            acc[mz][my][2] = mfma64(xr[mz] + xi[mz], yr + yi, acc[mz][my][2]);
          }
        }
      }
    }
    __syncthreads();
  }
  // raw sums: in place when the item has one chunk, else into the chunk's partial slot
#pragma unroll
  for (int mz = 0; mz < 2; ++mz)
#pragma unroll
    for (int my = 0; my < 2; ++my)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int zi = z0 + wz + 16 * mz + 4 * r + lk;
        const int yj = y0 + wy + 16 * my + li;
        if (zi >= nz || yj >= ncol) continue;
        const int64_t q = (int64_t)zi * ncol + yj;
        double* dst = it.nchunks == 1 ? it.out + 2 * q : it.part + 2 * ((int64_t)u.chunk * it.nq + q);
        if (GAUSS) {
          const double p1 = acc[mz][my][0][r], p2 = acc[mz][my][1][r], p3 = acc[mz][my][2][r];
          dst[0] = p1 - p2;
          dst[1] = p3 - p1 - p2;
        } else {
          dst[0] = acc[mz][my][0][r];
          dst[1] = acc[mz][my][1][r];
        }
      }
}

int product_rotate_records64(Context* c, const Mesh& m, const Rot& R, double* rec_out, hipStream_t s) {
  k_rotate_records64<<<ceil_div(m.ntri, 256), 256, 0, s>>>(m.d_rec64, m.ntri, R, rec_out);
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

void product64_launch(hipStream_t s, const ProductItem* items, const ProductUnit* units, int nunits,
                      const ProductItem& single, int nzt, int nyt, bool cq, bool gauss) {
  if (nunits <= 0) return;
  if (cq) {
    if (gauss) k_ff_product64<true, true><<<nunits, 256, 0, s>>>(items, units, nunits, single, nzt, nyt);
    else k_ff_product64<true, false><<<nunits, 256, 0, s>>>(items, units, nunits, single, nzt, nyt);
  } else {
    if (gauss) k_ff_product64<false, true><<<nunits, 256, 0, s>>>(items, units, nunits, single, nzt, nyt);
    else k_ff_product64<false, false><<<nunits, 256, 0, s>>>(items, units, nunits, single, nzt, nyt);
  }
}

}  // namespace hgx
