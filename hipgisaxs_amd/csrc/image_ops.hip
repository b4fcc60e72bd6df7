// image_ops.hip -- element-wise accumulation of form factors and intensities, the separable
// Gaussian smearing of a detector image, and the distance of an image to reference data (the host
// loop around the hot path, run on the device so the image never leaves HBM between grains,
// structures, the final smearing and a fit's objective).
#include <cmath>
#include <algorithm>
#include <cfloat>
#include <vector>

#include "hgx_internal.hpp"

namespace hgx {

__global__ __launch_bounds__(256) void k_caxpy(int64_t n, double ar, double ai, const double* __restrict__ x,
                                               bool accumulate, double* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double xr = x[2 * i], xi = x[2 * i + 1];
  double re = ar * xr - ai * xi, im = ar * xi + ai * xr;
  if (accumulate) {
    re += y[2 * i];
    im += y[2 * i + 1];
  }
  y[2 * i] = re;
  y[2 * i + 1] = im;
}

__global__ __launch_bounds__(256) void k_axpy(int64_t n, double a, const double* __restrict__ x, bool accumulate,
                                              double* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double v = a * x[i];
  y[i] = accumulate ? y[i] + v : v;
}

// y = a x^2 (+ y), elementwise; x may be y (no restrict)
__global__ __launch_bounds__(256) void k_axpy_sq(int64_t n, double a, const double* x, bool accumulate, double* y) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double xi = x[i];
  const double v = a * xi * xi;
  y[i] = accumulate ? y[i] + v : v;
}

// One 1-D pass of the smearing: out[line, i] = sum_k w_k in[line, k] / sum_k w_k over the window
// around i, for every line of the image. (stride_i, stride_line) select rows or columns, so one
// kernel serves both passes; consecutive threads take consecutive elements of the contiguous
// axis (i along a row, the line index along a column) so every load is coalesced.
// compat = 1 keeps the reference's window [i-h, i+h) with the weight indexed from the window
// start (include/numerics/convolutions.hpp:141-147), which shifts the kernel at the edges;
// compat = 0 weights by distance over the symmetric window [i-h, i+h].
__global__ __launch_bounds__(256) void k_smear_pass(int n, int nlines, int64_t stride_i, int64_t stride_line,
                                                    bool lines_fast, int h, const double* __restrict__ g,
                                                    int compat, const double* __restrict__ in,
                                                    double* __restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)n * nlines) return;
  int i, line;
  if (lines_fast) {
    line = (int)(t % nlines);
    i = (int)(t / nlines);
  } else {
    i = (int)(t % n);
    line = (int)(t / n);
  }
  const double* src = in + line * stride_line;
  int beg = max(0, i - h);
  int end = compat ? min(n, i + h) : min(n - 1, i + h) + 1;
  double sum = 0.0, norm = 0.0;
  for (int k = beg; k < end; ++k) {
    double w = g[compat ? k - beg : k - i + h];
    sum += src[k * stride_i] * w;
    norm += w;
  }
  out[line * stride_line + (int64_t)i * stride_i] = norm > DBL_MIN ? sum / norm : 0.0;
}

}  // namespace hgx

using namespace hgx;

extern "C" int hgx_caxpy(hgx_ctx* h, int64_t n, double a_re, double a_im, const double* x, int accumulate,
                         double* y) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (n < 0 || (n > 0 && (!x || !y))) return set_error(c, HGX_ERR_INVALID_ARG, "caxpy: bad arguments");
  if (n == 0) return HGX_OK;
  if (int rc = set_device(c)) return rc;
  LaunchTimer lt(c, false);
  k_caxpy<<<ceil_div(n, 256), 256, 0, c->stream>>>(n, a_re, a_im, x, accumulate != 0, y);
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

extern "C" int hgx_axpy(hgx_ctx* h, int64_t n, double a, const double* x, int accumulate, double* y) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (n < 0 || (n > 0 && (!x || !y))) return set_error(c, HGX_ERR_INVALID_ARG, "axpy: bad arguments");
  if (n == 0) return HGX_OK;
  if (int rc = set_device(c)) return rc;
  LaunchTimer lt(c, false);
  k_axpy<<<ceil_div(n, 256), 256, 0, c->stream>>>(n, a, x, accumulate != 0, y);
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

extern "C" int hgx_axpy_sq(hgx_ctx* h, int64_t n, double a, const double* x, int accumulate, double* y) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (n < 0 || (n > 0 && (!x || !y))) return set_error(c, HGX_ERR_INVALID_ARG, "axpy_sq: bad arguments");
  if (n == 0) return HGX_OK;
  if (int rc = set_device(c)) return rc;
  LaunchTimer lt(c, false);
  k_axpy_sq<<<ceil_div(n, 256), 256, 0, c->stream>>>(n, a, x, accumulate != 0, y);
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

// ----- distance of a model image to reference data ------------------------------------------------

namespace hgx {

// Pixel transform of the normalised functors: square root or cube root (ROOT 2 or 3).
template <int ROOT>
__device__ inline double root_of(double x) {
  return ROOT == 2 ? sqrt(x) : cbrt(x);
}

__device__ inline double nan0(double x) { return isnan(x) ? 0.0 : x; }

constexpr int kDistWG = 256;

// Sums [nb][K] of one block: the K values of each thread reduced over the block in a fixed tree.
template <int K>
__device__ inline void block_store(double (&v)[K], double* __restrict__ out) {
  __shared__ double red[K][kDistWG];
#pragma unroll
  for (int k = 0; k < K; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = kDistWG / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int k = 0; k < K; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) out[(int64_t)blockIdx.x * K + k] = red[k][0];
  }
}

// Pass 1 of the normalised functors: per block, norm_l2^2 of the transformed model image, vec_dot of
// the transformed images and norm_l2^2 of the transformed data, NaN terms dropped
// (include/analyzer/distance_functions.hpp:40-57).
template <int ROOT>
__global__ __launch_bounds__(kDistWG) void k_distance_stats(int64_t n, const double* __restrict__ I,
                                                            const double* __restrict__ O,
                                                            const double* __restrict__ m,
                                                            double* __restrict__ part) {
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t p = (int64_t)blockIdx.x * kDistWG + threadIdx.x; p < n; p += (int64_t)gridDim.x * kDistWG) {
    const double w = m ? m[p] : 1.0;
    const double u = root_of<ROOT>(I[p]), r = root_of<ROOT>(O[p]);
    v[0] += w * nan0(u * u);
    v[1] += w * nan0(fabs(u * r));
    v[2] += w * nan0(r * r);
  }
  block_store<3>(v, part);
}

// Sums the [nb][K] block partials in a fixed order; for the normalised functors turns them into the
// constants of pass 2: (dat_norm, ref_norm, c = dot / dat_norm^2).
template <int K>
__global__ __launch_bounds__(kDistWG) void k_distance_finish(int nb, bool consts, const double* __restrict__ part,
                                                             double* __restrict__ out) {
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kDistWG) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += part[(int64_t)b * K + k];
  }
  __shared__ double red[K][kDistWG];
#pragma unroll
  for (int k = 0; k < K; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = kDistWG / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int k = 0; k < K; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (consts) {
      const double ns = sqrt(red[0][0]), nr = sqrt(red[K - 1][0]);
      out[0] = ns;
      out[1] = nr;
      out[2] = red[1][0] / (ns * ns);
    } else {
      out[0] = red[0][0];
    }
  }
}

// Pass 2: the per-pixel terms of each functor summed per block. chi^2: m (I - O)^2 / max(O, 1);
// AbsoluteDifferenceSquare (:125-143): (m |O - I|)^2; unit-vector forms (:423-456, :543-575):
// (m |r/|r| - u/|u||)^2; c-norm forms (:460-499, :579-616): (m |r - c u|)^2, with u and r the
// transformed model and data. The square-root forms drop non-finite terms; L2, chi^2 and the cube-root
// forms keep them (a NaN pixel makes the distance NaN), as the reference's loops do.
template <int METRIC>
__global__ __launch_bounds__(kDistWG) void k_distance_terms(int64_t n, const double* __restrict__ I,
                                                            const double* __restrict__ O,
                                                            const double* __restrict__ m,
                                                            const double* __restrict__ cst,
                                                            double* __restrict__ part) {
  double v[1] = {0.0};
  double ns = 1.0, nr = 1.0, cc = 0.0;
  if (METRIC >= HGX_METRIC_SQRT_UNIT_L2) {
    ns = cst[0];
    nr = cst[1];
    cc = cst[2];
  }
  for (int64_t p = (int64_t)blockIdx.x * kDistWG + threadIdx.x; p < n; p += (int64_t)gridDim.x * kDistWG) {
    const double w = m ? m[p] : 1.0;
    const double a = I[p], b = O[p];
    if (METRIC == HGX_METRIC_CHI2) {
      v[0] += w * (a - b) * (a - b) / fmax(b, 1.0);
    } else if (METRIC == HGX_METRIC_L2) {
      const double t = w * fabs(b - a);
      v[0] += t * t;
    } else if (METRIC == HGX_METRIC_SQRT_UNIT_L2 || METRIC == HGX_METRIC_CBRT_UNIT_L2) {
      constexpr int R = METRIC == HGX_METRIC_SQRT_UNIT_L2 ? 2 : 3;
      double t = w * fabs(root_of<R>(b) / nr - root_of<R>(a) / ns);
      if (R == 2) t = isfinite(t) ? t : 0.0;
      v[0] += t * t;
    } else {
      constexpr int R = METRIC == HGX_METRIC_SQRT_C_L2 ? 2 : 3;
      double t = w * fabs(root_of<R>(b) - cc * root_of<R>(a));
      if (R == 2) t = isfinite(t) ? t : 0.0;
      v[0] += t * t;
    }
  }
  block_store<1>(v, part);
}

template <int METRIC>
void launch_terms(hipStream_t s, unsigned nb, int64_t n, const double* I, const double* O, const double* m,
                  const double* cst, double* part) {
  k_distance_terms<METRIC><<<nb, kDistWG, 0, s>>>(n, I, O, m, cst, part);
}

// ----- derivative of the distance along tangent images -----
//
// With u = I^(1/R), r = O^(1/R) (R = 2 or 3), the stats' sums v0 = sum m u^2, v1 = sum m |u r| and
// ns = sqrt v0, nr = sqrt(sum m r^2):
//   unit forms: D = sum m^2 e^2, e = r/nr - u/ns, and
//               dD/du_p = -2 m^2 e/ns + 2 E m u/ns^3 with E = sum m^2 e u;
//   c forms:    D = sum m^2 (r - c u)^2, c = v1/v0, and
//               dD/du_p = -2 c m^2 (r - c u) - 2 F m (|r| - 2 c u)/v0 with F = sum m^2 (r - c u) u;
// then dD/dI = dD/du u/(R I) (zero where I <= 0 or the pixel's term is dropped); chi^2 and L2 as in
// metric_seed.
template <int METRIC>
__device__ inline void unit_or_c(double a, double b, double w, const double* cst, double& u, double& r, double& t,
                                 bool& keep) {
  constexpr int R = (METRIC == HGX_METRIC_SQRT_UNIT_L2 || METRIC == HGX_METRIC_SQRT_C_L2) ? 2 : 3;
  u = root_of<R>(a);
  r = root_of<R>(b);
  if (METRIC == HGX_METRIC_SQRT_UNIT_L2 || METRIC == HGX_METRIC_CBRT_UNIT_L2)
    t = r / cst[1] - u / cst[0];
  else
    t = r - cst[2] * u;
  keep = R == 3 || isfinite(w * fabs(t));
}

// Pass 3: E (unit forms) or F (c forms) per block.
template <int METRIC>
__global__ __launch_bounds__(kDistWG) void k_distance_aux(int64_t n, const double* __restrict__ I,
                                                          const double* __restrict__ O, const double* __restrict__ m,
                                                          const double* __restrict__ cst, double* __restrict__ part) {
  double v[1] = {0.0};
  for (int64_t p = (int64_t)blockIdx.x * kDistWG + threadIdx.x; p < n; p += (int64_t)gridDim.x * kDistWG) {
    const double w = m ? m[p] : 1.0;
    double u, r, t;
    bool keep;
    unit_or_c<METRIC>(I[p], O[p], w, cst, u, r, t, keep);
    if (keep) v[0] += w * w * t * u;
  }
  block_store<1>(v, part);
}

// Pass 4: the seed g_p = dD/dI_p.
template <int METRIC>
__global__ __launch_bounds__(kDistWG) void k_distance_seed(int64_t n, const double* __restrict__ I,
                                                           const double* __restrict__ O, const double* __restrict__ m,
                                                           const double* __restrict__ cst, double* __restrict__ seed) {
  const int64_t p = (int64_t)blockIdx.x * kDistWG + threadIdx.x;
  if (p >= n) return;
  const double w = m ? m[p] : 1.0;
  const double a = I[p], b = O[p];
  double g = 0.0;
  if (METRIC == HGX_METRIC_CHI2) {
    g = 2.0 * w * (a - b) / fmax(b, 1.0);
  } else if (METRIC == HGX_METRIC_L2) {
    g = 2.0 * w * w * (a - b);
  } else {
    constexpr int R = (METRIC == HGX_METRIC_SQRT_UNIT_L2 || METRIC == HGX_METRIC_SQRT_C_L2) ? 2 : 3;
    double u, r, t;
    bool keep;
    unit_or_c<METRIC>(a, b, w, cst, u, r, t, keep);
    const double ns = cst[0];
    if (keep && a > 0.0 && ns > 0.0) {
      double du;
      if (METRIC == HGX_METRIC_SQRT_UNIT_L2 || METRIC == HGX_METRIC_CBRT_UNIT_L2)
        du = -2.0 * w * w * t / ns + 2.0 * cst[4] * w * u / (ns * ns * ns);
      else
        du = -2.0 * cst[2] * w * w * t - 2.0 * cst[4] * w * (fabs(r) - 2.0 * cst[2] * u) / (ns * ns);
      g = du * u / (R * a);
    }
  }
  seed[p] = g;
}

// Per block and direction j = blockIdx.y: sum_p seed_p dI[j][p].
__global__ __launch_bounds__(kDistWG) void k_seed_dots(int64_t n, const double* __restrict__ seed,
                                                       const double* __restrict__ dI, double* __restrict__ part) {
  const double* x = dI + (int64_t)blockIdx.y * n;
  double v[1] = {0.0};
  for (int64_t p = (int64_t)blockIdx.x * kDistWG + threadIdx.x; p < n; p += (int64_t)gridDim.x * kDistWG)
    v[0] += seed[p] * x[p];
  __shared__ double red[kDistWG];
  red[threadIdx.x] = v[0];
  __syncthreads();
  for (int w = kDistWG / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// out[j] = sum_b part[j][b], one workgroup per direction, fixed order.
__global__ __launch_bounds__(kDistWG) void k_dots_finish(int nb, const double* __restrict__ part,
                                                         double* __restrict__ out) {
  const double* x = part + (int64_t)blockIdx.x * nb;
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += kDistWG) v += x[b];
  __shared__ double red[kDistWG];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = kDistWG / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

}  // namespace hgx

extern "C" int hgx_gaussian_smear(hgx_ctx* h, int ncol, int nrow, double sigma, int flags, double* I) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (ncol < 1 || nrow < 1 || !I || !(sigma > 0.0) || !std::isfinite(sigma))
    return set_error(c, HGX_ERR_INVALID_ARG, "gaussian_smear: bad arguments");
  if (sigma > 1e6) return set_error(c, HGX_ERR_INVALID_ARG, "gaussian_smear: sigma %g too large", sigma);
  if (int rc = set_device(c)) return rc;
  // the kernel table of the reference, gaussian(i - h, sigma) for i in [0, 2h]
  // (include/numerics/convolutions.hpp:121-135)
  const int hw = (int)std::ceil(6 * sigma);
  const int gn = 2 * hw + 1;
  std::vector<double> g(gn);
  const double kPi = 3.14159265358979323846;
  for (int i = 0; i < gn; ++i) {
    int x = i - hw;
    g[i] = 1.0 / (sigma * std::sqrt(2 * kPi) * std::exp((x * x) / (2 * sigma * sigma)));
  }
  const int64_t npix = (int64_t)ncol * nrow;
  size_t tmp_b = ((npix * sizeof(double)) + 255) / 256 * 256;
  if (int rc = ensure_work(c, tmp_b + (gn * sizeof(double) + 255) / 256 * 256)) return rc;
  double* tmp = static_cast<double*>(c->d_work);
  double* dg = reinterpret_cast<double*>(static_cast<char*>(c->d_work) + tmp_b);
  {
    // through the pinned staging buffer: the host table may go out of scope at once
    const void* src[1] = {g.data()};
    const size_t nbytes[1] = {gn * sizeof(double)};
    if (int rc = stage_to_device(c, 1, src, nbytes, reinterpret_cast<char*>(dg))) return rc;
  }
  const int compat = (flags & HGX_SMEAR_COMPAT) ? 1 : 0;
  {
    LaunchTimer lt(c, false);
    // horizontal: along each row (i = column, contiguous)
    k_smear_pass<<<ceil_div(npix, 256), 256, 0, c->stream>>>(ncol, nrow, 1, ncol, false, hw, dg, compat, I, tmp);
    HGX_CHECK_LAUNCH(c);
  }
  {
    LaunchTimer lt(c, false);
    // vertical: along each column (i = row); threads run along the contiguous column index
    k_smear_pass<<<ceil_div(npix, 256), 256, 0, c->stream>>>(nrow, ncol, ncol, 1, true, hw, dg, compat, tmp, I);
    HGX_CHECK_LAUNCH(c);
  }
  return HGX_OK;
}

namespace hgx {

// Enqueues the distance passes; the value lands in device memory at *d_out. Workspace layout:
// [block partials][constants: ns, nr, c, D, E|F][seed n][direction partials].
static int distance_enqueue(Context* c, int64_t n, int metric, const double* I, const double* O, const double* mask,
                            double* d_out, int ndir, const double* dI, double* d_dout) {
  if (n < 1 || !I || !O) return set_error(c, HGX_ERR_INVALID_ARG, "image_distance: bad arguments");
  if (metric < HGX_METRIC_CHI2 || metric > HGX_METRIC_CBRT_C_L2)
    return set_error(c, HGX_ERR_INVALID_ARG, "image_distance: unknown metric %d", metric);
  if (ndir < 0 || (ndir > 0 && (!dI || !d_dout)) || ndir > 4096)
    return set_error(c, HGX_ERR_INVALID_ARG, "image_distance: bad tangent arguments");
  if (int rc = set_device(c)) return rc;
  // a fixed grid per size keeps the summation order, and so the value, reproducible
  const unsigned nb = (unsigned)std::min<int64_t>((n + kDistWG - 1) / kDistWG, 1024);
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t part_b = al((size_t)nb * 3 * sizeof(double)), cst_b = 256;
  const size_t seed_b = ndir > 0 ? al((size_t)n * sizeof(double)) : 0;
  const size_t dpart_b = ndir > 0 ? al((size_t)ndir * nb * sizeof(double)) : 0;
  if (int rc = ensure_work(c, part_b + cst_b + seed_b + dpart_b)) return rc;
  char* w = static_cast<char*>(c->d_work);
  double* part = reinterpret_cast<double*>(w);
  double* cst = reinterpret_cast<double*>(w + part_b);
  double* seed = reinterpret_cast<double*>(w + part_b + cst_b);
  double* dpart = reinterpret_cast<double*>(w + part_b + cst_b + seed_b);
  hipStream_t s = c->stream;
  const bool norm = metric >= HGX_METRIC_SQRT_UNIT_L2;
  if (norm) {
    const bool cube = metric == HGX_METRIC_CBRT_UNIT_L2 || metric == HGX_METRIC_CBRT_C_L2;
    {
      LaunchTimer lt(c, false);
      if (cube) k_distance_stats<3><<<nb, kDistWG, 0, s>>>(n, I, O, mask, part);
      else k_distance_stats<2><<<nb, kDistWG, 0, s>>>(n, I, O, mask, part);
      HGX_CHECK_LAUNCH(c);
    }
    k_distance_finish<3><<<1, kDistWG, 0, s>>>((int)nb, true, part, cst);
    HGX_CHECK_LAUNCH(c);
  }
  {
    LaunchTimer lt(c, false);
    switch (metric) {
      case HGX_METRIC_CHI2: launch_terms<HGX_METRIC_CHI2>(s, nb, n, I, O, mask, cst, part); break;
      case HGX_METRIC_L2: launch_terms<HGX_METRIC_L2>(s, nb, n, I, O, mask, cst, part); break;
      case HGX_METRIC_SQRT_UNIT_L2: launch_terms<HGX_METRIC_SQRT_UNIT_L2>(s, nb, n, I, O, mask, cst, part); break;
      case HGX_METRIC_SQRT_C_L2: launch_terms<HGX_METRIC_SQRT_C_L2>(s, nb, n, I, O, mask, cst, part); break;
      case HGX_METRIC_CBRT_UNIT_L2: launch_terms<HGX_METRIC_CBRT_UNIT_L2>(s, nb, n, I, O, mask, cst, part); break;
      default: launch_terms<HGX_METRIC_CBRT_C_L2>(s, nb, n, I, O, mask, cst, part); break;
    }
    HGX_CHECK_LAUNCH(c);
  }
  k_distance_finish<1><<<1, kDistWG, 0, s>>>((int)nb, false, part, cst + 3);
  HGX_CHECK_LAUNCH(c);
  HGX_CHECK_HIP(c, hipMemcpyAsync(d_out, cst + 3, sizeof(double), hipMemcpyDeviceToDevice, s));
  if (ndir == 0) return HGX_OK;
  LaunchTimer lt(c, false);
#define HGX_SEED(M)                                                                                \
  do {                                                                                             \
    if (M >= HGX_METRIC_SQRT_UNIT_L2) {                                                            \
      k_distance_aux<M><<<nb, kDistWG, 0, s>>>(n, I, O, mask, cst, part);                          \
      k_distance_finish<1><<<1, kDistWG, 0, s>>>((int)nb, false, part, cst + 4);                   \
    }                                                                                              \
    k_distance_seed<M><<<ceil_div(n, kDistWG), kDistWG, 0, s>>>(n, I, O, mask, cst, seed);         \
  } while (0)
  switch (metric) {
    case HGX_METRIC_CHI2: HGX_SEED(HGX_METRIC_CHI2); break;
    case HGX_METRIC_L2: HGX_SEED(HGX_METRIC_L2); break;
    case HGX_METRIC_SQRT_UNIT_L2: HGX_SEED(HGX_METRIC_SQRT_UNIT_L2); break;
    case HGX_METRIC_SQRT_C_L2: HGX_SEED(HGX_METRIC_SQRT_C_L2); break;
    case HGX_METRIC_CBRT_UNIT_L2: HGX_SEED(HGX_METRIC_CBRT_UNIT_L2); break;
    default: HGX_SEED(HGX_METRIC_CBRT_C_L2); break;
  }
#undef HGX_SEED
  HGX_CHECK_LAUNCH(c);
  k_seed_dots<<<dim3(nb, ndir), kDistWG, 0, s>>>(n, seed, dI, dpart);
  HGX_CHECK_LAUNCH(c);
  k_dots_finish<<<ndir, kDistWG, 0, s>>>((int)nb, dpart, d_dout);
  HGX_CHECK_LAUNCH(c);
  return HGX_OK;
}

}  // namespace hgx

extern "C" int hgx_image_distance(hgx_ctx* h, int64_t n, int metric, const double* I, const double* O,
                                  const double* mask, double* distance) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!distance) return set_error(c, HGX_ERR_INVALID_ARG, "image_distance: bad arguments");
  if (int rc = set_device(c)) return rc;
  if (int rc = ensure_result(c, 1)) return rc;
  if (int rc = distance_enqueue(c, n, metric, I, O, mask, c->d_result, 0, nullptr, nullptr)) return rc;
  HGX_CHECK_HIP(c, hipMemcpyAsync(distance, c->d_result, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return HGX_OK;
}

extern "C" int hgx_image_distance_async(hgx_ctx* h, int64_t n, int metric, const double* I, const double* O,
                                        const double* mask, double* d_distance) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!d_distance) return set_error(c, HGX_ERR_INVALID_ARG, "image_distance_async: null output");
  return distance_enqueue(c, n, metric, I, O, mask, d_distance, 0, nullptr, nullptr);
}

extern "C" int hgx_image_distance_tangent(hgx_ctx* h, int64_t n, int metric, const double* I, const double* O,
                                          const double* mask, int ndir, const double* dI, double* distance,
                                          double* ddistance) {
  Context* c = reinterpret_cast<Context*>(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!distance || ndir < 1 || !ddistance)
    return set_error(c, HGX_ERR_INVALID_ARG, "image_distance_tangent: bad arguments");
  if (int rc = set_device(c)) return rc;
  if (int rc = ensure_result(c, 1 + ndir)) return rc;
  if (int rc = distance_enqueue(c, n, metric, I, O, mask, c->d_result, ndir, dI, c->d_result + 1)) return rc;
  std::vector<double> host(1 + ndir);
  HGX_CHECK_HIP(c, hipMemcpyAsync(host.data(), c->d_result, host.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  *distance = host[0];
  for (int j = 0; j < ndir; ++j) ddistance[j] = host[1 + j];
  return HGX_OK;
}
