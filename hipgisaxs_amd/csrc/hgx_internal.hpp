// hgx_internal.hpp -- context state, error plumbing and device helpers shared by the hgx kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "hgx.h"

namespace hgx {

// One uploaded triangle mesh. Triangles are stored in a spatially sorted order (Morton order of
// the centroids) and grouped in blocks of kBlockTri; every block carries its own FP64 origin so the
// FP32 path only ever sees coordinates relative to a nearby point.
constexpr int kBlockTri = 128;
constexpr int kSeriesOrder = 8;  // small-|q| moment series: terms k = 0..kSeriesOrder
// number of moment components: sum_k 3 * (k+1)(k+2)/2
constexpr int kMomentCount = 3 * 165;

// Size of the per-device scratch array of the allocation-free entry point (hgx_ff_numeric_noalloc).
constexpr size_t kScratchBytes = (size_t)32 << 20;

// Process-wide serial numbers: each q-grid generation and each uploaded mesh takes a new one, so
// work prepared against a grid or a mesh (batches, grating plans) can tell that it was replaced.
uint64_t next_serial();

struct Mesh {
  bool live = false;
  uint64_t serial = 0;    // next_serial() at upload
  int kind = 0;
  int64_t ntri = 0;       // triangles (after dropping degenerate ones)
  int64_t ntri_in = 0;    // triangles as uploaded
  int64_t nblocks = 0;    // ceil(ntri / kBlockTri)
  double volume = 0.0;    // sum_t s_t (n_t . c_t) / 3
  double origin[3] = {0, 0, 0};
  double radius = 0.0;    // max_t |c_t - origin|
  double xradius = 0.0;   // max |v - origin| over the vertices (exact meshes)
  double snmax = 0.0;     // max_t |s_t n_t|
  // device arrays
  double* d_rec64 = nullptr;   // [ntri][8]: s*nx, s*ny, s*nz, cx, cy, cz, 0, 0 (absolute, FP64)
  float* d_rec32 = nullptr;    // [nblocks*kBlockTri][8]: s*n (3), c - block origin (3), 0, 0 (FP32)
  double* d_borig = nullptr;   // [nblocks][4]: block origin x, y, z, 0
  double* d_tri9 = nullptr;    // [ntri][9] vertices (exact meshes only), same order as d_rec64
  float* d_dv32 = nullptr;     // [nblocks*kBlockTri][8] vertices 1, 2 minus the centroid (exact meshes), FP32
  float* d_bdmax = nullptr;    // [nblocks] largest vertex-to-centroid distance in the block (exact meshes)
  double* d_moments = nullptr; // [kMomentCount] moments about origin (see ff_numeric.hip)
  double* d_xmoments = nullptr; // [kMomentCount] the same series for flat triangles (exact meshes)
  int64_t* d_src = nullptr;    // [ntri] caller's index of each stored (sorted) triangle
};

// A chi^2 plan: model constants and the device inputs it reads (caller-owned).
struct FitPlan {
  bool live = false;
  int metric = 0;  // HGX_METRIC_*
  hgx_fit_model model;
  const double* fc = nullptr;
  const double* obs = nullptr;
  const double* mask = nullptr;
};

// A prepared batch of numeric form factors (batch.hip): device work tables, partial-sum buffer
// and the captured hipGraph.
struct BatchState {
  bool live = false;
  int precision = 0, flags = 0, qpt = 4;
  bool cq = false;
  int nunits = 0, nfinish = 0;
  void* d_items = nullptr;
  void* d_units = nullptr;
  void* d_finish = nullptr;
  double* d_part = nullptr;
  // product-grid items (FP32C on product grids): their own unit list and rotated records
  int pmode = 0;
  int npunits = 0;
  void* d_pitems = nullptr;
  void* d_punits = nullptr;
  void* d_prot = nullptr;
  // Ewald-grid items (FP32C on QGrid::create / detector grids): units grouped by series order
  void* d_eitems = nullptr;
  void* d_eunits = nullptr;
  int neunits[4] = {0, 0, 0, 0};
  int eunit_off[4] = {0, 0, 0, 0};
  bool epipe = false;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  double pairs = 0.0;
  // what the work tables point into, with the serial each had at creation (checked before a run)
  std::vector<std::pair<const struct Context*, uint64_t>> grids;
  std::vector<std::pair<int, uint64_t>> meshes;
};

void free_batch_state(BatchState& b);

struct GratingPlan;  // grating.hip
struct Context;
void free_grating_plans(Context* c);

// ----- product-grid numeric form factor on the matrix cores (product_mfma.hip) ------------------
//
// A q-grid is a product grid when q(row i, column j, branch b) = (qx0, qy_j, qz_{b,i}): the Born
// q-space grid with qx = 0 and its four DWBA branches are. On such a grid the approximate-triangle
// sum factorizes per triangle, exp(i q.c) = exp(i(qx0 cx + qy cy)) exp(i qz cz), so the sum over
// triangles is a complex matrix product (z-rows x triangles) . (triangles x y-columns).

// One form factor for the product-grid kernel. Triangle records and block origins are already
// rotated (R^T applied), so the kernel works with the unrotated grid axes.
struct ProductItem {
  const double* qx;     // per-pixel grid arrays; qx[0] is qx0, qy[0..ncol) is the qy axis,
  const double* qy;     // qz row zi (zi = branch * nrow + row) is at qz + 2 * zi * ncol
  const double* qz;
  const float* rec;     // [nblocks * kBlockTri][8] rotated s*n (3), c - o_b (3), 0, 0
  const double* borig;  // [nblocks][4] rotated block origins
  const double* rec64;  // [ntri][8] rotated FP64 records s*n (3), c (3), 0, 0 (FP64 kernel only)
  int64_t ntri;         // triangles of rec64
  int64_t nblocks;
  int64_t bchunk;       // blocks per chunk
  int64_t nq;
  int nchunks;
  int ncol;
  int nz;               // nbranch * nrow
  float snmax;          // max_t |s_t n_t| over the mesh (split-fp16 operand scaling)
  float rmax;           // max_t |c_t| over the mesh (bounds exp(-Im(qz) c_z) for complex qz)
  int pad;
  double* part;         // [nchunks][nq] complex partial sums (nchunks > 1)
  double* out;          // [nq] complex raw sums (nchunks == 1)
  float* part32 = nullptr;  // FP32C kernels: FP32 partial sums here instead of part (exact: FP32 sums times 2^k)
};

// A work unit: one 32 (z) x 32 (y) tile of one item over one chunk of triangle blocks.
struct ProductUnit {
  int item;
  int ztile;
  int ytile;
  int chunk;
};

// Node dimensions of an analytic shape (HGX_SHAPE_* order); a truncated sphere adds its heights.
inline int analytic_dims(int shape, int flags) {
  static const int k[7] = {1, 2, 3, 1, 2, 2, 4};
  if (shape == HGX_SHAPE_SPHERE && (flags & HGX_SHAPE_TRUNCATED_SPHERE)) return 2;
  return k[shape];
}
// A truncated sphere's heights must not exceed the diameter of any of its radii (the slice radii
// sqrt(R^2 - z^2) are real); true when the nodes are usable.
inline bool analytic_heights_ok(int shape, int flags, const int n[4], const double* const val[4]) {
  if (shape != HGX_SHAPE_SPHERE || !(flags & HGX_SHAPE_TRUNCATED_SPHERE)) return true;
  for (int i = 0; i < n[0]; ++i)
    for (int k = 0; k < n[1]; ++k)
      if (!(val[1][k] >= 0.0 && val[1][k] <= 2.0 * val[0][i])) return false;
  return true;
}

// Planes of the split-fp16 adjoint weights: the real and imaginary parts, each as high and low fp16.
constexpr int kAdjW16Planes = 4;

// One mesh for the product-grid adjoint kernel (adjoint.hip).
struct AdjItem {
  const double* qx;     // grid arrays: qx[0] = qx0, qy[0..ncol) the qy axis, qz row z at qz + 2 z ncol
  const double* qy;
  const double* qz;
  const float* wt;      // [nq] complex Wt = -i W / |q|^2
  const double* rec;    // [ntri][8] rotated FP64 records (A', c')
  const float* rec32;   // [nblocks * kBlockTri][8] rotated FP32 records (A', c' - o_b)
  const double* borig;  // [nblocks][4] rotated block origins
  double* part;         // [6 forms][nyt][ntri] complex: B11, B1y, B1yy, Bz1, Bzy, Bzz per y-tile
  const uint4* w16;     // split-fp16 weights: [kAdjW16Planes = re, im x hi, lo][nzo][ncol] of 8 fp16 (one z octet, one column)
  const int* wexp;      // [ncol] power-of-two column scales of w16
  int64_t ntri;
  int nblocks;
  int ncol, nz, nyt;
  int nzo;              // ceil(nz / 8)
};

constexpr int kProductTile = 32;
constexpr int kProductWaves = 4;  // work units (waves) per workgroup
// product kernel variants: FP32 operands with 4 or 3 (Gauss) real products per complex product,
// or each FP32 operand split into two fp16 halves (4 real products, 3 cross terms each) with 1, 2
// or 4 row tiles of 32 per wave (two-level FP32 accumulation, or in place over the chunk: Flat)
constexpr int kProductF32x4 = 0;
constexpr int kProductF32Gauss = 1;
constexpr int kProductF16x2 = 2;
constexpr int kProductF16x2Z2 = 3;
constexpr int kProductF16x2Z4 = 4;
constexpr int kProductF16x2Z4Flat = 5;
// the same with the z operands built once per workgroup and shared through LDS: one work unit per
// workgroup, 32 x row tiles rows by 4 x 32 columns
constexpr int kProductF16x2S2 = 6;
constexpr int kProductF16x2S4Flat = 7;
constexpr int kProductF16x2S4 = 8;
// plain FP64 on the FP64 matrix cores (product_fp64.hip): 64 x 64 per workgroup, four or three
// (Gauss) real products per complex product
constexpr int kProductF64x4 = 9;
constexpr int kProductF64Gauss = 10;
inline bool product_is_f64(int mode) { return mode == kProductF64x4 || mode == kProductF64Gauss; }
inline bool product_is_f16(int mode) { return mode >= kProductF16x2 && mode <= kProductF16x2S4; }
// one work unit per workgroup
inline bool product_is_shared(int mode) { return mode >= kProductF16x2S2; }
// rows (z) and columns (y) of one work unit's tile for a kernel variant
inline int product_zrows(int mode) {
  return product_is_f64(mode) ? 2 * kProductTile
         : (mode == kProductF16x2Z2 || mode == kProductF16x2S2) ? 2 * kProductTile
         : (mode == kProductF16x2Z4 || mode == kProductF16x2Z4Flat || mode == kProductF16x2S4Flat ||
            mode == kProductF16x2S4)
             ? 4 * kProductTile
                                                                                 : kProductTile;
}
inline int product_ycols(int mode) {
  return product_is_f64(mode) ? 2 * kProductTile : product_is_shared(mode) ? kProductWaves * kProductTile : kProductTile;
}


// ----- Ewald-sphere q-grids on the matrix cores (ewald_mfma.hip) --------------------------------
//
// QGrid::create (src/model/qgrid.cpp:76-84) and the detector geometry give
//     q(row i, column j, branch b) = (k0 (cos af_i cos tth_j - cos ai), k0 cos af_i sin tth_j, qz_{b,i}),
// which is q = (a_i + c_i vx_j, c_i vy_j, qz_{b,i}) with a_i = k0 (cos af_i - cos ai), c_i = cos af_i
// and v_j = k0 (cos tth_j - 1, sin tth_j): a row part, and a small column vector scaled per row. Any
// grid of that form (fitted from the grid arrays, not assumed from how they were made) runs the
// approximate-triangle sum on the matrix cores: rows sorted by c are tiled 64 at a time, the tile's
// mid value cbar goes into the column operands exactly and exp(i (c_i - cbar) v_j . (c_t - o)) is
// expanded in a short series whose terms are products of row and column factors.
struct EwaldGrid {
  bool ok = false;
  int ntile = 0;            // 64-row tiles of the branch-major rows sorted by c
  double dmax = 0.0;        // largest half-range D of c over a tile
  double vmax = 0.0;        // largest |v_j|
  void* d_mem = nullptr;    // one allocation holding the four tables below
  const double* d_row = nullptr;   // [nrow][2] (a_i, c_i)
  const double* d_col = nullptr;   // [ncol][2] (vx_j, vy_j)
  const double* d_tile = nullptr;  // [ntile][2] (cbar, D)
  const int* d_perm = nullptr;     // [ntile * 64] branch-major row of each tile slot, -1 past the end
};

// One form factor for the Ewald-grid kernel; records and block origins rotated by R^T as for the
// product kernel, o = R^T (mesh origin) the expansion point.
struct EwaldItem {
  const double* qz;     // grid qz: row z = branch * nrow + row at qz + 2 z ncol
  const double* row;
  const double* col;
  const double* tile;
  const int* perm;
  const float* rec;     // [nblocks * kBlockTri][8] rotated s*n (3), c - o_b (3), 0, 0
  const double* borig;  // [nblocks][4] rotated block origins
  double ox, oy, oz;
  int64_t nblocks;
  int64_t bchunk;
  int64_t nq;
  int nchunks;
  int ncol, nrow, ntile;
  float snmax;          // max_t |s_t n_t|
  float rmax;           // bound on |c_t| (complex qz attenuation)
  double* part;         // [nchunks][nq] complex partial sums (nchunks > 1)
  double* out;          // [nq] complex raw sums (nchunks == 1)
  float* part32 = nullptr;  // FP32C kernels: FP32 partial sums here instead of part (exact: FP32 sums times 2^k)
};

// Launch-shape settings of the numeric kernels, per context (hgx_tune_*; the autotuner sets them).
// Batches and grating plans copy what they need when they are created.
struct Tuning {
  int qpt64 = 2, qpt32 = 4, target_wg = 2048;  // per-pair kernels: q-points per thread, workgroup target
  int batch_qpt = 4, batch_bpu = 16;            // batched per-pair kernels: q-points, blocks per unit
  int product_bpu = 64;                         // product-grid kernels: triangle blocks per unit
  int product_operands = HGX_PRODUCT_F16X2;
  int product_mults = 4;
  int product_mz = 2, product_two = 1, product_shared = 1;
  int adjoint_variant = 4;  // 0: FP32 operands; 2 or 4: split fp16 with 2 or 4 column tiles per wave
  int product64_mults = 4;  // FP64 product-grid kernel: 4 or 3 (Gauss) real products per complex product
  int ewald_pipe = 0;       // Ewald-grid kernel: column operands built a step ahead (1) or in their step (0)
};

struct Context {
  int device = 0;
  Tuning tune;
  std::vector<FitPlan> fit_plans;
  std::vector<BatchState> batches;
  std::vector<GratingPlan*> gratings;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  // q-grid
  int ncol = 0, nrow = 0, nbranch = 0;
  int64_t npix = 0, nq = 0;
  bool qz_real = true;      // all Im(qz) == 0
  bool product = false;     // q(row, col, branch) = (qx0, qy[col], qz[branch, row]) exactly
  bool qz_rows = false;     // qz depends on (branch, row) only
  EwaldGrid ew;             // the Ewald-grid factorization, when the grid has one
  int* d_flag = nullptr;    // device scratch flags of the grid checks (2 ints)
  bool have_alpha = false;
  double* d_qx = nullptr;   // [npix]
  double* d_qy = nullptr;   // [npix]
  double* d_qz = nullptr;   // [2*nq] interleaved complex
  double* d_alpha = nullptr;// [nrow]
  uint64_t grid_gen = 0;    // next_serial() whenever the grid arrays are replaced or freed
  // meshes
  std::vector<Mesh> meshes;
  // scratch for partial sums
  void* d_work = nullptr;
  size_t work_bytes = 0;
  double* d_result = nullptr;  // small results copied back to the host (distances and their tangents)
  int64_t result_cap = 0;
  // device address of the device's static scratch array (kScratchBytes), shared by every context on
  // the device; null when it could not be resolved
  char* d_scratch = nullptr;
  // pinned host staging of small per-call tables (stage_to_device), reused once its last copy is done
  void* h_stage = nullptr;
  size_t stage_bytes = 0;
  hipEvent_t stage_done = nullptr;
  std::string last_error;
  // optional HIP-event timing of kernel launches (hgx_profile_begin / hgx_profile_end)
  bool profiling = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_main, ev_aux, ev_pool;
};

// Brackets one kernel launch with events when profiling is on. `main` marks the hot kernel of
// the call (the triangle / node loop), aux everything else (finish, epilogues).
struct LaunchTimer {
  Context* c;
  bool main;
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  LaunchTimer(Context* ctx, bool is_main);
  ~LaunchTimer();
};

// ----- error plumbing ----------------------------------------------------------------------

int set_error(Context* c, int code, const char* fmt, ...);

#define HGX_CHECK_HIP(ctx, expr)                                                              \
  do {                                                                                        \
    hipError_t e__ = (expr);                                                                  \
    if (e__ != hipSuccess)                                                                    \
      return ::hgx::set_error((ctx), HGX_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr,          \
                              hipGetErrorString(e__), __FILE__, __LINE__);                     \
  } while (0)

// launches are followed by hipGetLastError so a bad launch is reported on the call that made it
#define HGX_CHECK_LAUNCH(ctx) HGX_CHECK_HIP(ctx, hipGetLastError())

int ensure_work(Context* c, size_t bytes);
// d_result with room for n doubles (grows; synchronises the stream before it replaces the buffer).
int ensure_result(Context* c, int64_t n);
// Copies n host tables (sizes in bytes) into device memory dst, each at a 256-byte aligned offset,
// through the context's pinned staging buffer and one asynchronous copy on the context's stream; the
// caller's tables may be released on return (no stream synchronisation).
int stage_to_device(Context* c, int n, const void* const* src, const size_t* bytes, char* dst);
int set_device(Context* c);

// ----- device helpers ------------------------------------------------------------------------

struct cd {
  double r, i;
};

__host__ __device__ inline cd mk(double r, double i) { return cd{r, i}; }
__host__ __device__ inline cd operator+(cd a, cd b) { return cd{a.r + b.r, a.i + b.i}; }
__host__ __device__ inline cd operator-(cd a, cd b) { return cd{a.r - b.r, a.i - b.i}; }
__host__ __device__ inline cd operator-(cd a) { return cd{-a.r, -a.i}; }
__host__ __device__ inline cd operator*(cd a, cd b) { return cd{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__host__ __device__ inline cd operator*(double s, cd a) { return cd{s * a.r, s * a.i}; }
__host__ __device__ inline cd operator*(cd a, double s) { return cd{s * a.r, s * a.i}; }
__host__ __device__ inline double norm2(cd a) { return a.r * a.r + a.i * a.i; }
__host__ __device__ inline cd conj(cd a) { return cd{a.r, -a.i}; }
__device__ inline cd cdiv(cd a, cd b) {
  // Smith's algorithm keeps the quotient finite for widely scaled operands
  if (fabs(b.r) >= fabs(b.i)) {
    double t = b.i / b.r, d = b.r + b.i * t;
    return cd{(a.r + a.i * t) / d, (a.i - a.r * t) / d};
  }
  double t = b.r / b.i, d = b.r * t + b.i;
  return cd{(a.r * t + a.i) / d, (a.i * t - a.r) / d};
}
__device__ inline cd mul_i(cd a) { return cd{-a.i, a.r}; }      // i * a
__device__ inline cd mul_mi(cd a) { return cd{a.i, -a.r}; }     // -i * a
// The complex functions below take a shortcut for an exactly real argument (a particle in a
// non-absorbing layer makes every q real): exp(-0) = cosh(0) = 1 and sinh(+-0) = +-0, sin(+-0) = +-0
// exactly, so the shortcut returns the same bits as the general formula while skipping its
// exponentials.
__device__ inline cd cexpi_d(cd t) {                             // exp(i t)
  double s, c;
  sincos(t.r, &s, &c);
  if (t.i == 0.0) return cd{c, s};
  double m = exp(-t.i);
  return cd{m * c, m * s};
}
__device__ inline cd cexp_d(cd z) {
  double m = exp(z.r);
  if (z.i == 0.0) return cd{m, m * z.i};
  double s, c;
  sincos(z.i, &s, &c);
  return cd{m * c, m * s};
}
__device__ inline cd csqrt_d(cd z) {
  // principal branch, as std::sqrt(std::complex)
  double a = hypot(z.r, z.i);
  if (a == 0.0) return cd{0.0, 0.0};
  if (z.r >= 0.0) {
    double t = sqrt(0.5 * (a + z.r));
    return cd{t, z.i / (2.0 * t)};
  }
  double t = sqrt(0.5 * (a - z.r));
  return cd{fabs(z.i) / (2.0 * t), copysign(t, z.i)};
}
__device__ inline cd csin_d(cd z) {  // sin(a+ib) = sin a cosh b + i cos a sinh b
  double s, c;
  sincos(z.r, &s, &c);
  if (z.i == 0.0) return cd{s, c * z.i};
  return cd{s * cosh(z.i), c * sinh(z.i)};
}
__device__ inline cd ccos_d(cd z) {  // cos(a+ib) = cos a cosh b - i sin a sinh b
  double s, c;
  sincos(z.r, &s, &c);
  if (z.i == 0.0) return cd{c, -s * z.i};
  return cd{c * cosh(z.i), -s * sinh(z.i)};
}
__device__ inline double cabs_d(cd z) { return z.i == 0.0 ? fabs(z.r) : hypot(z.r, z.i); }

// Rotation of a q-point: mq = R (qx, qy, qz) with complex qz (matrix.hpp:123-130).
struct Rot {
  double m[9];
};
__device__ inline void rotate_q(const Rot& R, double qx, double qy, cd qz, cd mq[3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    mq[j].r = R.m[3 * j] * qx + R.m[3 * j + 1] * qy + R.m[3 * j + 2] * qz.r;
    mq[j].i = R.m[3 * j + 2] * qz.i;
  }
}

struct Vec3 {
  double v[3];
};

inline Rot make_rot(const double* r) {
  Rot R;
  for (int i = 0; i < 9; ++i) R.m[i] = r ? r[i] : ((i % 4 == 0) ? 1.0 : 0.0);
  return R;
}
inline Vec3 make_vec(const double* v) {
  Vec3 o;
  for (int i = 0; i < 3; ++i) o.v[i] = v ? v[i] : 0.0;
  return o;
}

// The square-root normalised distance functors from six masked sums of the model image I and the
// data O: S = (sum m I, sum m sqrt(IO), sum m O, sum m^2 I, sum m^2 sqrt(IO), sum m^2 O) (norm_l2 and
// vec_dot weight by the mask, the squared distance by its square; include/analyzer/
// distance_functions.hpp:40-57, 423-499). Returns the distance D and the constants k[0..2] of the
// per-pixel seed dD/dI, with s = sqrt I, r = sqrt O:
//   sqrt-unit-L2: D = B3/A3 - 2 B2/(Nr Ns) + B1/A1 (Ns^2 = A1, Nr^2 = A3), L = B1/Ns - B2/Nr,
//                 dD/dI = m^2 k0 - m^2 k1 r/s - m k2 with k = (1/A1, 1/(Nr Ns), L/Ns^3);
//   sqrt-c-L2:    c = A2/A1, K = B2 - c B1, D = B3 - 2 c B2 + c^2 B1,
//                 dD/dI = -m^2 k0 (r/s - k0) - m k1 (r/s - 2 k0) with k = (c, K/A1).
// A zero model image (A1 = 0) is taken as u = 0 (sqrt-unit) and c = 0.
__host__ __device__ inline double metric_from_sums(int metric, const double* S, double k[3]) {
  const double A1 = S[0], A2 = S[1], A3 = S[2], B1 = S[3], B2 = S[4], B3 = S[5];
  k[0] = k[1] = k[2] = 0.0;
  if (metric == HGX_METRIC_SQRT_UNIT_L2) {
    const double Nr = sqrt(A3), Ns = sqrt(A1);
    double D = Nr > 0 ? B3 / A3 : 0.0;
    if (Ns > 0) {
      D += B1 / A1 - (Nr > 0 ? 2.0 * B2 / (Nr * Ns) : 0.0);
      const double L = B1 / Ns - (Nr > 0 ? B2 / Nr : 0.0);
      k[0] = 1.0 / A1;
      k[1] = Nr > 0 ? 1.0 / (Nr * Ns) : 0.0;
      k[2] = L / (A1 * Ns);
    }
    return D;
  }
  const double cc = A1 > 0 ? A2 / A1 : 0.0;
  const double K = B2 - cc * B1;
  k[0] = cc;
  k[1] = A1 > 0 ? K / A1 : 0.0;
  return B3 - 2.0 * cc * B2 + cc * cc * B1;
}

// dD/dI of one pixel for every metric (chi^2, L2 and the two normalised forms with the constants
// above); term is the pixel's own share of D (0 for the normalised forms, whose D comes from the sums).
__host__ __device__ inline double metric_seed(int metric, double I, double O, double m, const double k[3],
                                              double* term) {
  if (metric == HGX_METRIC_CHI2) {
    const double w = m / fmax(O, 1.0);
    *term = w * (I - O) * (I - O);
    return 2.0 * w * (I - O);
  }
  if (metric == HGX_METRIC_L2) {
    const double d = m * (I - O);
    *term = d * d;
    return 2.0 * m * d;
  }
  *term = 0.0;
  const double sI = I > 0.0 ? sqrt(I) : 0.0, sO = O > 0.0 ? sqrt(O) : 0.0;
  if (!(sI > 0.0)) return 0.0;
  const double ratio = sO / sI;
  return metric == HGX_METRIC_SQRT_UNIT_L2 ? m * m * k[0] - m * m * k[1] * ratio - m * k[2]
                                           : -m * m * k[0] * (ratio - k[0]) - k[1] * m * (ratio - 2.0 * k[0]);
}

inline unsigned ceil_div(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// Rotates an item's block-relative FP32 records and FP64 block origins by R^T.
int product_rotate_mesh(Context* c, const Mesh& m, const Rot& R, float* rec_out, double* borig_out,
                        hipStream_t s);
// Launches the unit list of a batch (units sorted by the caller).
void product_launch_units(hipStream_t s, const ProductItem* d_items, const ProductUnit* d_units, int nunits,
                          bool cq, int mode);
// Launches every (z-tile, y-tile, chunk) unit of a single item.
void product_launch_single(hipStream_t s, const ProductItem& item, bool cq, int mode);
// Launches the product-grid adjoint over items (d_items == NULL: the single item by value);
// upi workgroup units per item (3 f-passes x triangle blocks x y-tiles).
// Split-fp16 form of the adjoint weights for nitems items (strides in elements): per-column maxima
// into pmax ([nitems][ceil(nz / 256)][ncol] floats), then w16 and wexp as AdjItem describes.
void adjoint_split_weights(hipStream_t s, int nitems, const float* wt, int64_t wt_stride, int nz, int ncol,
                           float* pmax, int64_t pmax_stride, uint4* w16, int64_t w16_stride, int* wexp,
                           int64_t wexp_stride);
int64_t adjoint_pmax_count(int nz, int ncol);  // floats of pmax per item
// work units per item of the product-grid adjoint for the current operand setting
// adjoint kernel variant: 0 FP32 operands, 2 or 4 split fp16 with that many column tiles per wave
int64_t adjoint_product_units(const AdjItem& it, int variant);
int adjoint_product_variant(const Context* c);  // the context's setting (hgx_tune_adjoint)
void adjoint_product_launch(hipStream_t s, const AdjItem* d_items, const AdjItem& single, int upi, int nunits,
                            bool cq, int variant);
// Tuning: triangle blocks per work unit and the kernel variant (kProduct* above); the FP64 variant.
// A work unit of a batch's Ewald-grid items: one 64-row tile x 128-column group of one item over one
// chunk of triangle blocks.
struct EwaldUnit {
  int item;
  int tile;
  int yg;
  int chunk;
};

// Ewald-grid tables of the current grid (ewald_mfma.hip): fitted after every grid change, freed with
// the grid. ewald_orders gives the series orders K (0..3) the matrix-core path needs for a mesh of
// this radius about its origin, or -1 when it would need more (the per-pair kernels then run).
int ewald_update(Context* c);
void ewald_free(Context* c);
int ewald_orders(const Context* c, double radius);
int ewald_launch(Context* c, const EwaldItem& it, bool cq, int orders);
// The grid, mesh and rotation fields of an item (expansion origin R^T o, tables, sizes); the caller
// sets rec, borig, bchunk, nchunks, part and out.
void ewald_fill_item(const Context* grid, const Mesh& m, const Rot& R, EwaldItem& it);
// Column groups (128 columns) of a grid of ncol columns.
inline int ewald_col_groups(int ncol) { return (ncol + kProductWaves * kProductTile - 1) / (kProductWaves * kProductTile); }
// A batch's units of one series order (d_units), in one launch.
void ewald_launch_units(hipStream_t s, const EwaldItem* d_items, const EwaldUnit* d_units, int nunits, bool cq,
                        int orders, bool pipelined);
int product_blocks_per_unit(const Context* c);
int product_mode(const Context* c);
inline int product_mode64(const Context* c);
// FP64 product-grid kernel (product_fp64.hip): rotated FP64 records (R^T applied) and the launch
int product_rotate_records64(Context* c, const Mesh& m, const Rot& R, double* rec_out, hipStream_t s);
void product64_launch(hipStream_t s, const ProductItem* items, const ProductUnit* units, int nunits,
                      const ProductItem& single, int nzt, int nyt, bool cq, bool gauss);

inline int product_mode64(const Context* c) { return c->tune.product64_mults == 3 ? kProductF64Gauss : kProductF64x4; }

}  // namespace hgx
