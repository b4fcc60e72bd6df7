// image_ops.hip -- element-wise accumulation of form factors and intensities, and the separable
// Gaussian smearing of a detector image (the host loop around the hot path, run on the device so
// the image never leaves HBM between grains, structures and the final smearing).
#include <cmath>
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
  if (int rc = ensure_work(c, tmp_b + gn * sizeof(double))) return rc;
  double* tmp = static_cast<double*>(c->d_work);
  double* dg = reinterpret_cast<double*>(static_cast<char*>(c->d_work) + tmp_b);
  HGX_CHECK_HIP(c, hipMemcpyAsync(dg, g.data(), gn * sizeof(double), hipMemcpyHostToDevice, c->stream));
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
  // the host table must outlive the asynchronous copy
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return HGX_OK;
}
