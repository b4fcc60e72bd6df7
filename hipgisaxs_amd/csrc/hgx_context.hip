// hgx_context.hip -- context lifetime, q-grid construction on the device, mesh upload.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <numeric>

#include "hgx_internal.hpp"

namespace hgx {

int set_error(Context* c, int code, const char* fmt, ...) {
  if (!c) return code;
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  c->last_error = buf;
  return code;
}

int set_device(Context* c) {
  HGX_CHECK_HIP(c, hipSetDevice(c->device));
  return HGX_OK;
}

int ensure_work(Context* c, size_t bytes) {
  if (bytes <= c->work_bytes) return HGX_OK;
  if (c->d_work) {
    // the stream may still read the old buffer
    HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    HGX_CHECK_HIP(c, hipFree(c->d_work));
    c->d_work = nullptr;
    c->work_bytes = 0;
  }
  size_t sz = std::max<size_t>(bytes, (size_t)1 << 20);
  if (hipMalloc(&c->d_work, sz) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(c, HGX_ERR_OOM, "workspace allocation of %zu bytes failed", sz);
  }
  c->work_bytes = sz;
  return HGX_OK;
}

int ensure_result(Context* c, int64_t n) {
  if (n <= c->result_cap) return HGX_OK;
  if (c->d_result) {
    HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
    HGX_CHECK_HIP(c, hipFree(c->d_result));
    c->d_result = nullptr;
    c->result_cap = 0;
  }
  const int64_t cap = std::max<int64_t>(n, 4096);
  if (hipMalloc(&c->d_result, (size_t)cap * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(c, HGX_ERR_OOM, "result buffer allocation failed");
  }
  c->result_cap = cap;
  return HGX_OK;
}

int stage_to_device(Context* c, int n, const void* const* src, const size_t* bytes, char* dst) {
  size_t total = 0;
  for (int i = 0; i < n; ++i) total += (bytes[i] + 255) / 256 * 256;
  if (total == 0) return HGX_OK;
  if (!c->stage_done) HGX_CHECK_HIP(c, hipEventCreateWithFlags(&c->stage_done, hipEventDisableTiming));
  // the previous staged copy must have read the buffer before it is rewritten
  HGX_CHECK_HIP(c, hipEventSynchronize(c->stage_done));
  if (total > c->stage_bytes) {
    if (c->h_stage) HGX_CHECK_HIP(c, hipHostFree(c->h_stage));
    c->h_stage = nullptr;
    c->stage_bytes = 0;
    const size_t sz = std::max<size_t>(total, (size_t)1 << 16);
    if (hipHostMalloc(&c->h_stage, sz, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      c->h_stage = nullptr;
      return set_error(c, HGX_ERR_OOM, "pinned staging allocation of %zu bytes failed", sz);
    }
    c->stage_bytes = sz;
  }
  char* h = static_cast<char*>(c->h_stage);
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    if (bytes[i]) std::memcpy(h + off, src[i], bytes[i]);
    off += (bytes[i] + 255) / 256 * 256;
  }
  HGX_CHECK_HIP(c, hipMemcpyAsync(dst, h, total, hipMemcpyHostToDevice, c->stream));
  HGX_CHECK_HIP(c, hipEventRecord(c->stage_done, c->stream));
  return HGX_OK;
}

LaunchTimer::LaunchTimer(Context* ctx, bool is_main) : c(ctx), main(is_main) {
  if (!c->profiling) return;
  if (!c->ev_pool.empty()) {
    ev = c->ev_pool.back();
    c->ev_pool.pop_back();
  } else if (hipEventCreate(&ev.first) != hipSuccess || hipEventCreate(&ev.second) != hipSuccess) {
    (void)hipGetLastError();
    ev = {nullptr, nullptr};
    return;
  }
  (void)hipEventRecord(ev.first, c->stream);
}

LaunchTimer::~LaunchTimer() {
  if (!ev.first) return;
  (void)hipEventRecord(ev.second, c->stream);
  (main ? c->ev_main : c->ev_aux).push_back(ev);
}

static void free_mesh(Mesh& m) {
  if (m.d_rec64) (void)hipFree(m.d_rec64);
  if (m.d_rec32) (void)hipFree(m.d_rec32);
  if (m.d_borig) (void)hipFree(m.d_borig);
  if (m.d_tri9) (void)hipFree(m.d_tri9);
  if (m.d_moments) (void)hipFree(m.d_moments);
  if (m.d_src) (void)hipFree(m.d_src);
  if (m.d_dv32) (void)hipFree(m.d_dv32);
  if (m.d_xmoments) (void)hipFree(m.d_xmoments);
  if (m.d_bdmax) (void)hipFree(m.d_bdmax);
  m = Mesh();
}

// Scratch memory of the allocation-free entry point (hgx_ff_numeric_noalloc): a static array that
// each device holds until the process ends. Host code never uses the name `scratch` as an address:
// its device address is resolved once per device, with that device current, and kept.
static __device__ __align__(256) char scratch[kScratchBytes];

static char* device_scratch(int device) {
  static std::mutex mu;
  static std::vector<char*> ptrs;
  static std::vector<char> done;
  std::lock_guard<std::mutex> lock(mu);
  if ((int)ptrs.size() <= device) {
    ptrs.resize(device + 1, nullptr);
    done.resize(device + 1, 0);
  }
  if (!done[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) {
      ptrs[device] = static_cast<char*>(p);
      done[device] = 1;  // a failed lookup is tried again by the next context on the device
    } else {
      (void)hipGetLastError();
    }
  }
  return ptrs[device];
}

uint64_t next_serial() {
  static std::atomic<uint64_t> counter{0};
  return ++counter;
}

static void free_qgrid(Context* c) {
  c->grid_gen = next_serial();
  if (c->d_qx) (void)hipFree(c->d_qx);
  if (c->d_qy) (void)hipFree(c->d_qy);
  if (c->d_qz) (void)hipFree(c->d_qz);
  if (c->d_alpha) (void)hipFree(c->d_alpha);
  c->d_qx = c->d_qy = c->d_qz = c->d_alpha = nullptr;
  c->npix = c->nq = 0;
  c->ncol = c->nrow = c->nbranch = 0;
  c->have_alpha = false;
  c->product = false;
  c->qz_rows = false;
  ewald_free(c);
}

// Allocates grid arrays for nrow x ncol pixels and nbranch branches.
static int alloc_qgrid(Context* c, int ncol, int nrow, int nbranch) {
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  free_qgrid(c);
  int64_t npix = (int64_t)ncol * nrow;
  int64_t nq = npix * nbranch;
  if (hipMalloc(&c->d_qx, npix * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->d_qy, npix * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->d_qz, 2 * nq * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->d_alpha, nrow * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    free_qgrid(c);
    return set_error(c, HGX_ERR_OOM, "q-grid allocation failed (%lld pixels)", (long long)npix);
  }
  c->ncol = ncol;
  c->nrow = nrow;
  c->nbranch = nbranch;
  c->npix = npix;
  c->nq = nq;
  return HGX_OK;
}

// ----- q-grid kernels ------------------------------------------------------------------------

// q-space region of QGrid::create (src/model/qgrid.cpp:44-85). One thread per pixel; alpha_f of
// the unreversed row nrow-1-i and theta of column j are recomputed per pixel.
__global__ void k_qgrid_qspace(int ncol, int nrow, double qy_min, double qz_min, double qy_max, double qz_max,
                               double alpha_i, double k0, double* __restrict__ qx, double* __restrict__ qy,
                               double* __restrict__ qz, double* __restrict__ alpha_f) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)ncol * nrow) return;
  int i = (int)(p / ncol), j = (int)(p % ncol);
  double sin_ai = sin(alpha_i), cos_ai = cos(alpha_i);
  double dqz = (qz_max - qz_min) / (nrow - 1);
  int src_row = nrow - 1 - i;  // alpha_ is reversed after construction (:75)
  double alf = asin((qz_min + src_row * dqz) / k0 - sin_ai);
  double alpha0 = asin(qz_min / k0 - sin_ai);  // alpha_[0] before the reversal (:57)
  double cos_af = cos(alpha0);
  double dqy = (qy_max - qy_min) / (ncol - 1);
  double qpt = qy_min + j * dqy;
  double kf = qpt / k0;
  double tmp = (cos_af * cos_af + cos_ai * cos_ai - kf * kf) / (2 * cos_af * cos_ai);
  if (tmp > 1.0) tmp = 1.0;
  double tth = (qpt < 0 ? -1.0 : 1.0) * acos(tmp);
  qx[p] = k0 * (cos(alf) * cos(tth) - cos_ai);
  qy[p] = k0 * (cos(alf) * sin(tth));
  qz[2 * p] = k0 * (sin(alf) + sin_ai);
  qz[2 * p + 1] = 0.0;
  if (j == 0) alpha_f[i] = alf;
}

__global__ void k_qgrid_explicit(int ncol, int nrow, double qy_min, double qz_min, double qy_max, double qz_max,
                                 double alpha_i, double k0, double* __restrict__ qx, double* __restrict__ qy,
                                 double* __restrict__ qz, double* __restrict__ alpha_f) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)ncol * nrow) return;
  int i = (int)(p / ncol), j = (int)(p % ncol);
  double qzv = qz_min + (nrow - 1 - i) * ((qz_max - qz_min) / (nrow - 1));
  qx[p] = 0.0;
  qy[p] = qy_min + j * ((qy_max - qy_min) / (ncol - 1));
  qz[2 * p] = qzv;
  qz[2 * p + 1] = 0.0;
  if (j == 0) alpha_f[i] = asin(qzv / k0 - sin(alpha_i));
}

__global__ void k_qgrid_detector(int ncol, int nrow, double pixel, double distance, double col_center,
                                 double row_horizon, double alpha_i, double k0, double* __restrict__ qx,
                                 double* __restrict__ qy, double* __restrict__ qz, double* __restrict__ alpha_f) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)ncol * nrow) return;
  int i = (int)(p / ncol), j = (int)(p % ncol);
  double alf = atan(((nrow - 1 - i) + 0.5 - row_horizon) * pixel / distance);
  double tth = atan((j - col_center) * pixel / distance);
  qx[p] = k0 * (cos(alf) * cos(tth) - cos(alpha_i));
  qy[p] = k0 * (cos(alf) * sin(tth));
  qz[2 * p] = k0 * (sin(alf) + sin(alpha_i));
  qz[2 * p + 1] = 0.0;
  if (j == 0) alpha_f[i] = alf;
}

// QGrid::create_qz_extended (src/model/qgrid.cpp:102-127); reads branch-0 qz (real part) in place
// of the per-pixel qz_ and writes the four branches.
__global__ void k_qz_extended(int64_t npix, const double* __restrict__ qz_in, double k0, double alpha_i,
                              double dnl_re, double dnl_im, double* __restrict__ qz_out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  double sin_ai = sin(alpha_i);
  double kzi_0 = -k0 * sin_ai;
  cd kzi = -k0 * csqrt_d(cd{sin_ai * sin_ai - dnl_re, -dnl_im});
  double kzf_0 = qz_in[2 * i] + kzi_0;
  double sg = kzf_0 < 0 ? -1.0 : 1.0;
  cd kzf = sg * csqrt_d(cd{kzf_0 * kzf_0 - k0 * k0 * dnl_re, -k0 * k0 * dnl_im});
  cd b[4] = {kzf - kzi, -kzf - kzi, kzf + kzi, -kzf + kzi};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    qz_out[2 * (i + k * npix)] = b[k].r;
    qz_out[2 * (i + k * npix) + 1] = b[k].i;
  }
}

// flag[0] is cleared unless every q-point equals the product of the grid's axes: qx = qx[0], qy =
// qy[column] of row 0 and qz = qz of column 0 of its (branch, row); flag[1] unless qz depends on the
// (branch, row) only.
__global__ void k_check_product(int ncol, int64_t npix, int nbranch, const double* __restrict__ qx,
                                const double* __restrict__ qy, const double* __restrict__ qz, int* __restrict__ flag) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  int64_t row = p / ncol, j = p % ncol;
  bool rows = true;
  for (int b = 0; b < nbranch; ++b) {
    int64_t q = b * npix + p, q0 = b * npix + row * ncol;
    rows = rows && qz[2 * q] == qz[2 * q0] && qz[2 * q + 1] == qz[2 * q0 + 1];
  }
  if (!(rows && qx[p] == qx[0] && qy[p] == qy[j])) flag[0] = 0;  // every writer stores the same value
  if (!rows) flag[1] = 0;
}

}  // namespace hgx

using namespace hgx;

static Context* ctx_of(hgx_ctx* c) { return reinterpret_cast<Context*>(c); }

// Decides whether the current grid is a product grid and whether it has the Ewald-grid
// factorization (run once per grid change; synchronous).
static int update_product(Context* c) {
  c->product = false;
  c->qz_rows = false;
  ewald_free(c);
  if (!c->d_flag && hipMalloc(&c->d_flag, 2 * sizeof(int)) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(c, HGX_ERR_OOM, "q-grid: flag allocation failed");
  }
  int one[2] = {1, 1}, got[2] = {0, 0};
  HGX_CHECK_HIP(c, hipMemcpyAsync(c->d_flag, one, sizeof(one), hipMemcpyHostToDevice, c->stream));
  k_check_product<<<ceil_div(c->npix, 256), 256, 0, c->stream>>>(c->ncol, c->npix, c->nbranch, c->d_qx, c->d_qy,
                                                                 c->d_qz, c->d_flag);
  HGX_CHECK_LAUNCH(c);
  HGX_CHECK_HIP(c, hipMemcpyAsync(got, c->d_flag, sizeof(got), hipMemcpyDeviceToHost, c->stream));
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  c->product = got[0] != 0;
  c->qz_rows = got[1] != 0;
  return c->product || !c->qz_rows ? HGX_OK : ewald_update(c);
}
static const Context* ctx_of(const hgx_ctx* c) { return reinterpret_cast<const Context*>(c); }

extern "C" {

const char* hgx_version(void) { return "hgx 0.1 (gfx950)"; }

int hgx_create(int device, hgx_ctx** out) {
  if (!out) return HGX_ERR_INVALID_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    return HGX_ERR_DEVICE;
  }
  if (device < 0 || device >= ndev) return HGX_ERR_INVALID_ARG;
  Context* c = new (std::nothrow) Context();
  if (!c) return HGX_ERR_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return HGX_ERR_DEVICE;
  }
  c->stream = c->own_stream;
  c->d_scratch = device_scratch(device);  // with the device current
  *out = reinterpret_cast<hgx_ctx*>(c);
  return HGX_OK;
}

int hgx_destroy(hgx_ctx* h) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& m : c->meshes) free_mesh(m);
  for (auto& b : c->batches) free_batch_state(b);
  free_grating_plans(c);
  free_qgrid(c);
  for (auto* v : {&c->ev_main, &c->ev_aux, &c->ev_pool})
    for (auto& e : *v) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  if (c->d_work) (void)hipFree(c->d_work);
  if (c->d_result) (void)hipFree(c->d_result);
  if (c->d_flag) (void)hipFree(c->d_flag);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  if (c->stage_done) (void)hipEventDestroy(c->stage_done);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return HGX_OK;
}

int hgx_set_stream(hgx_ctx* h, void* stream) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  // NULL is the device's null stream (what PyTorch's default stream is), not "own stream"
  c->stream = reinterpret_cast<hipStream_t>(stream);
  return HGX_OK;
}

int hgx_use_own_stream(hgx_ctx* h) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  c->stream = c->own_stream;
  return HGX_OK;
}

int hgx_synchronize(hgx_ctx* h) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (int rc = set_device(c)) return rc;
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return HGX_OK;
}

int hgx_device_alloc(hgx_ctx* h, size_t bytes, void** ptr) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!ptr) return set_error(c, HGX_ERR_INVALID_ARG, "device_alloc: null output pointer");
  *ptr = nullptr;
  if (int rc = set_device(c)) return rc;
  if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return set_error(c, HGX_ERR_OOM, "device_alloc of %zu bytes failed", bytes);
  }
  return HGX_OK;
}

int hgx_device_free(hgx_ctx* h, void* ptr) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!ptr) return HGX_OK;
  if (int rc = set_device(c)) return rc;
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  HGX_CHECK_HIP(c, hipFree(ptr));
  return HGX_OK;
}

int hgx_memcpy_d2h(hgx_ctx* h, void* dst, const void* src, size_t bytes) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!dst || !src) return set_error(c, HGX_ERR_INVALID_ARG, "memcpy_d2h: null pointer");
  if (int rc = set_device(c)) return rc;
  HGX_CHECK_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return HGX_OK;
}

int hgx_memcpy_h2d(hgx_ctx* h, void* dst, const void* src, size_t bytes) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!dst || !src) return set_error(c, HGX_ERR_INVALID_ARG, "memcpy_h2d: null pointer");
  if (int rc = set_device(c)) return rc;
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  HGX_CHECK_HIP(c, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return HGX_OK;
}

int hgx_profile_begin(hgx_ctx* h) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  for (auto* v : {&c->ev_main, &c->ev_aux}) {
    for (auto& e : *v) c->ev_pool.push_back(e);
    v->clear();
  }
  c->profiling = true;
  return HGX_OK;
}

int hgx_profile_end(hgx_ctx* h, double* main_ms, int64_t* main_launches, double* aux_ms, int64_t* aux_launches) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  c->profiling = false;
  if (int rc = set_device(c)) return rc;
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  double ms[2] = {0.0, 0.0};
  int64_t n[2] = {0, 0};
  int k = 0;
  for (auto* v : {&c->ev_main, &c->ev_aux}) {
    for (auto& e : *v) {
      float t = 0.0f;
      HGX_CHECK_HIP(c, hipEventElapsedTime(&t, e.first, e.second));
      ms[k] += t;
      n[k] += 1;
      c->ev_pool.push_back(e);
    }
    v->clear();
    ++k;
  }
  if (main_ms) *main_ms = ms[0];
  if (main_launches) *main_launches = n[0];
  if (aux_ms) *aux_ms = ms[1];
  if (aux_launches) *aux_launches = n[1];
  return HGX_OK;
}

const char* hgx_last_error(const hgx_ctx* h) {
  const Context* c = ctx_of(h);
  if (!c) return "invalid context handle";
  return c->last_error.c_str();
}

int hgx_qgrid_qspace(hgx_ctx* h, int ncol, int nrow, double qy_min, double qz_min, double qy_max, double qz_max,
                     double alpha_i, double k0) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (ncol < 2 || nrow < 2 || !(k0 > 0)) return set_error(c, HGX_ERR_INVALID_ARG, "qgrid_qspace: need ncol, nrow >= 2 and k0 > 0");
  if (int rc = set_device(c)) return rc;
  if (int rc = alloc_qgrid(c, ncol, nrow, 1)) return rc;
  k_qgrid_qspace<<<ceil_div(c->npix, 256), 256, 0, c->stream>>>(ncol, nrow, qy_min, qz_min, qy_max, qz_max, alpha_i,
                                                               k0, c->d_qx, c->d_qy, c->d_qz, c->d_alpha);
  HGX_CHECK_LAUNCH(c);
  c->qz_real = true;
  c->have_alpha = true;
  return update_product(c);
}

int hgx_qgrid_explicit(hgx_ctx* h, int ncol, int nrow, double qy_min, double qz_min, double qy_max,
                       double qz_max, double alpha_i, double k0) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (ncol < 2 || nrow < 2 || !(k0 > 0)) return set_error(c, HGX_ERR_INVALID_ARG, "qgrid_explicit: need ncol, nrow >= 2 and k0 > 0");
  if (int rc = set_device(c)) return rc;
  if (int rc = alloc_qgrid(c, ncol, nrow, 1)) return rc;
  k_qgrid_explicit<<<ceil_div(c->npix, 256), 256, 0, c->stream>>>(ncol, nrow, qy_min, qz_min, qy_max, qz_max,
                                                                 alpha_i, k0, c->d_qx, c->d_qy, c->d_qz, c->d_alpha);
  HGX_CHECK_LAUNCH(c);
  c->qz_real = true;
  c->have_alpha = true;
  return update_product(c);
}

int hgx_qgrid_detector(hgx_ctx* h, int ncol, int nrow, double pixel, double distance, double col_center,
                       double row_horizon, double alpha_i, double k0) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (ncol < 1 || nrow < 1 || !(distance > 0) || !(pixel > 0) || !(k0 > 0))
    return set_error(c, HGX_ERR_INVALID_ARG, "qgrid_detector: bad geometry");
  if (int rc = set_device(c)) return rc;
  if (int rc = alloc_qgrid(c, ncol, nrow, 1)) return rc;
  k_qgrid_detector<<<ceil_div(c->npix, 256), 256, 0, c->stream>>>(ncol, nrow, pixel, distance, col_center,
                                                                 row_horizon, alpha_i, k0, c->d_qx, c->d_qy,
                                                                 c->d_qz, c->d_alpha);
  HGX_CHECK_LAUNCH(c);
  c->qz_real = true;
  c->have_alpha = true;
  return update_product(c);
}

int hgx_qgrid_set(hgx_ctx* h, int ncol, int nrow, int nbranch, const double* qx, const double* qy,
                  const double* qz_ext, const double* alpha_f) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (ncol < 1 || nrow < 1 || (nbranch != 1 && nbranch != 4) || !qx || !qy || !qz_ext)
    return set_error(c, HGX_ERR_INVALID_ARG, "qgrid_set: bad arguments");
  if (int rc = set_device(c)) return rc;
  if (int rc = alloc_qgrid(c, ncol, nrow, nbranch)) return rc;
  HGX_CHECK_HIP(c, hipMemcpy(c->d_qx, qx, c->npix * sizeof(double), hipMemcpyHostToDevice));
  HGX_CHECK_HIP(c, hipMemcpy(c->d_qy, qy, c->npix * sizeof(double), hipMemcpyHostToDevice));
  HGX_CHECK_HIP(c, hipMemcpy(c->d_qz, qz_ext, 2 * c->nq * sizeof(double), hipMemcpyHostToDevice));
  c->have_alpha = alpha_f != nullptr;
  if (alpha_f) HGX_CHECK_HIP(c, hipMemcpy(c->d_alpha, alpha_f, nrow * sizeof(double), hipMemcpyHostToDevice));
  bool real = true;
  for (int64_t i = 0; i < c->nq && real; ++i) real = qz_ext[2 * i + 1] == 0.0;
  c->qz_real = real;
  return update_product(c);
}

int hgx_qz_extended(hgx_ctx* h, double k0, double alpha_i, double dnl_re, double dnl_im) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "qz_extended: no q-grid");
  if (c->nbranch != 1) return set_error(c, HGX_ERR_INVALID_ARG, "qz_extended: grid already has %d branches", c->nbranch);
  if (int rc = set_device(c)) return rc;
  double* nq = nullptr;
  if (hipMalloc(&nq, 2 * 4 * c->npix * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(c, HGX_ERR_OOM, "qz_extended: allocation failed");
  }
  k_qz_extended<<<ceil_div(c->npix, 256), 256, 0, c->stream>>>(c->npix, c->d_qz, k0, alpha_i, dnl_re, dnl_im, nq);
  HGX_CHECK_LAUNCH(c);
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  HGX_CHECK_HIP(c, hipFree(c->d_qz));
  c->d_qz = nq;
  c->grid_gen = next_serial();
  c->nbranch = 4;
  c->nq = 4 * c->npix;
  // kzi and kzf are real exactly when the layer is lossless and every kzf0^2 >= k0^2 dnl
  c->qz_real = (dnl_re == 0.0 && dnl_im == 0.0);
  // the branches depend on the row only, so a product grid stays one; the Ewald-grid tables are
  // rebuilt for the four branches
  return update_product(c);
}

int hgx_qgrid_info(const hgx_ctx* h, int64_t* npix, int64_t* nq, int* nrow, int* ncol, int* nbranch) {
  const Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (npix) *npix = c->npix;
  if (nq) *nq = c->nq;
  if (nrow) *nrow = c->nrow;
  if (ncol) *ncol = c->ncol;
  if (nbranch) *nbranch = c->nbranch;
  return c->d_qz ? HGX_OK : HGX_ERR_NO_QGRID;
}

int hgx_qgrid_is_product(const hgx_ctx* h, int* is_product) {
  const Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (is_product) *is_product = c->product ? 1 : 0;
  return c->d_qz ? HGX_OK : HGX_ERR_NO_QGRID;
}

int hgx_qgrid_get(hgx_ctx* h, double* qx, double* qy, double* qz_ext, double* alpha_f) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!c->d_qz) return set_error(c, HGX_ERR_NO_QGRID, "qgrid_get: no q-grid");
  if (int rc = set_device(c)) return rc;
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  if (qx) HGX_CHECK_HIP(c, hipMemcpy(qx, c->d_qx, c->npix * sizeof(double), hipMemcpyDeviceToHost));
  if (qy) HGX_CHECK_HIP(c, hipMemcpy(qy, c->d_qy, c->npix * sizeof(double), hipMemcpyDeviceToHost));
  if (qz_ext) HGX_CHECK_HIP(c, hipMemcpy(qz_ext, c->d_qz, 2 * c->nq * sizeof(double), hipMemcpyDeviceToHost));
  if (alpha_f && c->have_alpha)
    HGX_CHECK_HIP(c, hipMemcpy(alpha_f, c->d_alpha, c->nrow * sizeof(double), hipMemcpyDeviceToHost));
  return HGX_OK;
}

// ----- meshes ----------------------------------------------------------------------------------

static uint64_t spread3(uint64_t x) {  // 21 bits -> every third bit
  x &= 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}

// Moment layout: for k = 0..K, component i = 0..2 of n, monomials (a,b,c) of degree k with a
// descending then b descending. M = sum_t s_t n_{t,i} (c_t - o)^(a,b,c).
static void accumulate_moments(const double* sn, const double* d, double* M) {
  double px[kSeriesOrder + 1], py[kSeriesOrder + 1], pz[kSeriesOrder + 1];
  px[0] = py[0] = pz[0] = 1.0;
  for (int k = 1; k <= kSeriesOrder; ++k) {
    px[k] = px[k - 1] * d[0];
    py[k] = py[k - 1] * d[1];
    pz[k] = pz[k - 1] * d[2];
  }
  int idx = 0;
  for (int k = 0; k <= kSeriesOrder; ++k)
    for (int i = 0; i < 3; ++i)
      for (int a = k; a >= 0; --a)
        for (int b = k - a; b >= 0; --b) {
          int cc = k - a - b;
          M[idx++] += sn[i] * px[a] * py[b] * pz[cc];
        }
}

// Index of monomial (a, b, n - a - b) among those of degree n in the moment layout's order.
static inline int mono_index(int n, int a, int b) { return (n - a) * (n - a + 1) / 2 + (n - a - b); }

// R = (w . q) P(q) for P homogeneous of degree n - 1, both as coefficients of the monomials of q.
static void mul_linear(const double* P, int n, const double* w, double* R) {
  for (int a = n; a >= 0; --a)
    for (int b = n - a; b >= 0; --b) {
      const int cc = n - a - b;
      double v = 0.0;
      if (a > 0) v += w[0] * P[mono_index(n - 1, a - 1, b)];
      if (b > 0) v += w[1] * P[mono_index(n - 1, a, b - 1)];
      if (cc > 0) v += w[2] * P[mono_index(n - 1, a, b)];
      R[mono_index(n, a, b)] = v;
    }
}

// The same moments for flat triangles with vertices o + w_j (exact meshes):
// M = sum_t s_t n_{t,i} <(r - o)^(a,b,c)>_t, the average over triangle t. It follows from
// <(q.(r - o))^n>_t = 2 n!/(n+2)! h_n(q.w_1, q.w_2, q.w_3), with h_n the complete homogeneous
// symmetric polynomial, built degree by degree from h_n(x1..x3) = h_n(x1, x2) + x3 h_{n-1}(x1..x3),
// h_n(x1, x2) = x1^n + x2 h_{n-1}(x1, x2). The symmetric-tensor entry of a monomial is its
// coefficient times a! b! c! / n!.
static void accumulate_moments_exact(const double* sn, const double (*w)[3], double* M) {
  constexpr int kTop = (kSeriesOrder + 1) * (kSeriesOrder + 2) / 2;
  double x1[kTop], h2[kTop], h3[kTop], t1[kTop], t2[kTop], t3[kTop];
  double fact[kSeriesOrder + 3];
  fact[0] = 1.0;
  for (int k = 1; k <= kSeriesOrder + 2; ++k) fact[k] = fact[k - 1] * k;
  x1[0] = h2[0] = h3[0] = 1.0;
  int idx = 0;
  for (int n = 0; n <= kSeriesOrder; ++n) {
    const int nm = (n + 1) * (n + 2) / 2;
    if (n > 0) {
      mul_linear(x1, n, w[0], t1);
      mul_linear(h2, n, w[1], t2);
      mul_linear(h3, n, w[2], t3);
      for (int j = 0; j < nm; ++j) {
        t2[j] += t1[j];
        t3[j] += t2[j];
      }
      std::copy(t1, t1 + nm, x1);
      std::copy(t2, t2 + nm, h2);
      std::copy(t3, t3 + nm, h3);
    }
    const double scale = 2.0 / fact[n + 2];
    for (int i = 0; i < 3; ++i)
      for (int a = n; a >= 0; --a)
        for (int b = n - a; b >= 0; --b) {
          const int cc = n - a - b;
          M[idx++] += sn[i] * scale * fact[a] * fact[b] * fact[cc] * h3[mono_index(n, a, b)];
        }
  }
}

int hgx_mesh_upload(hgx_ctx* h, int kind, int64_t ntri_in, const double* rec, int* mesh_id) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (!rec || !mesh_id || ntri_in < 1 || (kind != HGX_MESH_APPROX8 && kind != HGX_MESH_EXACT9))
    return set_error(c, HGX_ERR_INVALID_ARG, "mesh_upload: bad arguments");
  if (int rc = set_device(c)) return rc;
  // 1. approximate records (s, n, c) and, for exact meshes, vertices; drop degenerate triangles
  // the way the OBJ reader does (src/file/objectshape_reader.cpp:271-300).
  std::vector<double> s_n_c;  // 7 per triangle
  std::vector<double> tri9;
  std::vector<int64_t> kept;  // caller's index of every kept triangle
  s_n_c.reserve(7 * ntri_in);
  for (int64_t t = 0; t < ntri_in; ++t) {
    if (kind == HGX_MESH_APPROX8) {
      const double* r = rec + 8 * t;
      bool finite = true;
      for (int k = 0; k < 7; ++k) finite = finite && std::isfinite(r[k]);
      if (!finite) return set_error(c, HGX_ERR_INVALID_ARG, "mesh_upload: non-finite record %lld", (long long)t);
      if (r[0] == 0.0) continue;
      s_n_c.insert(s_n_c.end(), r, r + 7);
      kept.push_back(t);
    } else {
      const double* v = rec + 9 * t;
      for (int k = 0; k < 9; ++k)
        if (!std::isfinite(v[k])) return set_error(c, HGX_ERR_INVALID_ARG, "mesh_upload: non-finite vertex in %lld", (long long)t);
      double a[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
      double b[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
      double na = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      double nb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      double dt = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
      double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
      double nc = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
      if (nc == 0 || na == 0 || nb == 0) continue;
      double ct = dt / (na * nb);
      double area = na * nb * std::sqrt(1 - ct * ct) / 2;
      double r7[7] = {area, cr[0] / nc, cr[1] / nc, cr[2] / nc, (v[0] + v[3] + v[6]) / 3.0,
                      (v[1] + v[4] + v[7]) / 3.0, (v[2] + v[5] + v[8]) / 3.0};
      s_n_c.insert(s_n_c.end(), r7, r7 + 7);
      tri9.insert(tri9.end(), v, v + 9);
      kept.push_back(t);
    }
  }
  int64_t ntri = (int64_t)(s_n_c.size() / 7);
  if (ntri < 1) return set_error(c, HGX_ERR_INVALID_ARG, "mesh_upload: no non-degenerate triangles");
  // 2. Morton order of the centroids
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t t = 0; t < ntri; ++t)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], s_n_c[7 * t + 4 + k]);
      hi[k] = std::max(hi[k], s_n_c[7 * t + 4 + k]);
    }
  std::vector<uint64_t> key(ntri);
  for (int64_t t = 0; t < ntri; ++t) {
    uint64_t code = 0;
    for (int k = 0; k < 3; ++k) {
      double ext = hi[k] - lo[k];
      double u = ext > 0 ? (s_n_c[7 * t + 4 + k] - lo[k]) / ext : 0.0;
      uint64_t qv = (uint64_t)std::min(2097151.0, std::max(0.0, u * 2097151.0));
      code |= spread3(qv) << k;
    }
    key[t] = code;
  }
  std::vector<int64_t> order(ntri);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
  // 3. host images of the device arrays
  Mesh m;
  m.kind = kind;
  m.ntri = ntri;
  m.ntri_in = ntri_in;
  std::vector<int64_t> src(ntri);
  for (int64_t t = 0; t < ntri; ++t) src[t] = kept[order[t]];
  m.nblocks = (ntri + kBlockTri - 1) / kBlockTri;
  std::vector<double> rec64(8 * ntri, 0.0);
  std::vector<float> rec32(8 * m.nblocks * kBlockTri, 0.0f);
  std::vector<double> borig(4 * m.nblocks, 0.0);
  std::vector<double> tri9s(kind == HGX_MESH_EXACT9 ? 9 * ntri : 0);
  std::vector<float> dv32(kind == HGX_MESH_EXACT9 ? 8 * m.nblocks * kBlockTri : 0, 0.0f);
  std::vector<float> bdmax(kind == HGX_MESH_EXACT9 ? m.nblocks : 0, 0.0f);
  double vol = 0.0, orig[3] = {0, 0, 0};
  for (int64_t t = 0; t < ntri; ++t) {
    const double* r = &s_n_c[7 * order[t]];
    double* o = &rec64[8 * t];
    o[0] = r[0] * r[1];
    o[1] = r[0] * r[2];
    o[2] = r[0] * r[3];
    o[3] = r[4];
    o[4] = r[5];
    o[5] = r[6];
    vol += r[0] * (r[1] * r[4] + r[2] * r[5] + r[3] * r[6]);
    for (int k = 0; k < 3; ++k) orig[k] += r[4 + k];
    if (kind == HGX_MESH_EXACT9) std::copy(&tri9[9 * order[t]], &tri9[9 * order[t]] + 9, &tri9s[9 * t]);
  }
  m.volume = vol / 3.0;
  for (int k = 0; k < 3; ++k) m.origin[k] = orig[k] / ntri;
  double rad = 0.0;
  std::vector<double> M(kMomentCount, 0.0);
  for (int64_t t = 0; t < ntri; ++t) {
    const double* o = &rec64[8 * t];
    double d[3] = {o[3] - m.origin[0], o[4] - m.origin[1], o[5] - m.origin[2]};
    rad = std::max(rad, std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
    accumulate_moments(o, d, M.data());
  }
  m.radius = rad;
  std::vector<double> XM(kind == HGX_MESH_EXACT9 ? kMomentCount : 0, 0.0);
  if (kind == HGX_MESH_EXACT9) {
    double xr = 0.0;
    for (int64_t t = 0; t < ntri; ++t) {
      double w[3][3];
      for (int j = 0; j < 3; ++j) {
        for (int k = 0; k < 3; ++k) w[j][k] = tri9s[9 * t + 3 * j + k] - m.origin[k];
        xr = std::max(xr, std::sqrt(w[j][0] * w[j][0] + w[j][1] * w[j][1] + w[j][2] * w[j][2]));
      }
      accumulate_moments_exact(&rec64[8 * t], w, XM.data());
    }
    m.xradius = xr;
  }
  {
    // A closed surface has sum_t s_t n_t = 0 (divergence theorem); what the sum holds then is
    // rounding, which the series would divide by |q| and so blow up as q -> 0. Zero it for meshes
    // closed to rounding; open meshes keep their (physical) 1/|q| term.
    double asum = 0.0, amax = 0.0;
    for (int64_t t = 0; t < ntri; ++t) {
      double a = std::sqrt(rec64[8 * t] * rec64[8 * t] + rec64[8 * t + 1] * rec64[8 * t + 1] +
                           rec64[8 * t + 2] * rec64[8 * t + 2]);
      asum += a;
      amax = std::max(amax, a);
    }
    m.snmax = amax;
    double n0 = std::sqrt(M[0] * M[0] + M[1] * M[1] + M[2] * M[2]);
    if (n0 <= 1e-8 * asum) {
      M[0] = M[1] = M[2] = 0.0;
      if (!XM.empty()) XM[0] = XM[1] = XM[2] = 0.0;  // the same sums
    }
  }
  for (int64_t b = 0; b < m.nblocks; ++b) {
    int64_t t0 = b * kBlockTri, t1 = std::min<int64_t>(ntri, t0 + kBlockTri);
    double bo[3] = {0, 0, 0};
    for (int64_t t = t0; t < t1; ++t)
      for (int k = 0; k < 3; ++k) bo[k] += rec64[8 * t + 3 + k];
    for (int k = 0; k < 3; ++k) borig[4 * b + k] = bo[k] / (double)(t1 - t0);
    for (int64_t t = t0; t < t1; ++t) {
      const double* o = &rec64[8 * t];
      float* f = &rec32[8 * t];
      f[0] = (float)o[0];
      f[1] = (float)o[1];
      f[2] = (float)o[2];
      f[3] = (float)(o[3] - borig[4 * b]);
      f[4] = (float)(o[4] - borig[4 * b + 1]);
      f[5] = (float)(o[5] - borig[4 * b + 2]);
    }
    // padding triangles of the last block keep s*n = 0 and contribute exactly zero
    if (kind == HGX_MESH_EXACT9) {
      double dm = 0.0;
      for (int64_t t = t0; t < t1; ++t) {
        const double* v = &tri9s[9 * t];
        const double* o = &rec64[8 * t];
        for (int i = 0; i < 3; ++i) {
          double d2 = 0.0;
          for (int k = 0; k < 3; ++k) {
            const double d = v[3 * i + k] - o[3 + k];
            d2 += d * d;
            if (i < 2) dv32[8 * t + 3 * i + k] = (float)d;
          }
          dm = std::max(dm, std::sqrt(d2));
        }
      }
      bdmax[b] = (float)dm;
    }
  }
  // 4. device copies
  auto fail = [&](const char* what) {
    (void)hipGetLastError();
    free_mesh(m);
    return set_error(c, HGX_ERR_OOM, "mesh_upload: allocation of %s failed", what);
  };
  if (hipMalloc(&m.d_rec64, rec64.size() * sizeof(double)) != hipSuccess) return fail("rec64");
  if (hipMalloc(&m.d_rec32, rec32.size() * sizeof(float)) != hipSuccess) return fail("rec32");
  if (hipMalloc(&m.d_borig, borig.size() * sizeof(double)) != hipSuccess) return fail("block origins");
  if (hipMalloc(&m.d_moments, M.size() * sizeof(double)) != hipSuccess) return fail("moments");
  if (hipMalloc(&m.d_src, src.size() * sizeof(int64_t)) != hipSuccess) return fail("index map");
  if (kind == HGX_MESH_EXACT9 && hipMalloc(&m.d_tri9, tri9s.size() * sizeof(double)) != hipSuccess) return fail("tri9");
  if (kind == HGX_MESH_EXACT9 && hipMalloc(&m.d_dv32, dv32.size() * sizeof(float)) != hipSuccess)
    return fail("vertex offsets");
  if (kind == HGX_MESH_EXACT9 && hipMalloc(&m.d_bdmax, bdmax.size() * sizeof(float)) != hipSuccess)
    return fail("block extents");
  if (kind == HGX_MESH_EXACT9 && hipMalloc(&m.d_xmoments, XM.size() * sizeof(double)) != hipSuccess)
    return fail("triangle moments");
  auto h2d = [](void* dst, const void* src, size_t n) {
    return hipMemcpy(dst, src, n, hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!h2d(m.d_rec64, rec64.data(), rec64.size() * sizeof(double)) ||
      !h2d(m.d_rec32, rec32.data(), rec32.size() * sizeof(float)) ||
      !h2d(m.d_borig, borig.data(), borig.size() * sizeof(double)) ||
      !h2d(m.d_moments, M.data(), M.size() * sizeof(double)) ||
      !h2d(m.d_src, src.data(), src.size() * sizeof(int64_t)) ||
      (kind == HGX_MESH_EXACT9 && !h2d(m.d_tri9, tri9s.data(), tri9s.size() * sizeof(double))) ||
      (kind == HGX_MESH_EXACT9 && !h2d(m.d_dv32, dv32.data(), dv32.size() * sizeof(float))) ||
      (kind == HGX_MESH_EXACT9 && !h2d(m.d_bdmax, bdmax.data(), bdmax.size() * sizeof(float))) ||
      (kind == HGX_MESH_EXACT9 && !h2d(m.d_xmoments, XM.data(), XM.size() * sizeof(double)))) {
    hipError_t e = hipGetLastError();
    free_mesh(m);
    return set_error(c, HGX_ERR_DEVICE, "mesh_upload: copy to the device failed: %s", hipGetErrorString(e));
  }
  m.live = true;
  m.serial = next_serial();
  int id = -1;
  for (size_t i = 0; i < c->meshes.size(); ++i)
    if (!c->meshes[i].live) {
      id = (int)i;
      break;
    }
  if (id < 0) {
    id = (int)c->meshes.size();
    c->meshes.push_back(Mesh());
  }
  c->meshes[id] = m;
  *mesh_id = id;
  return HGX_OK;
}

int hgx_mesh_release(hgx_ctx* h, int mesh_id) {
  Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (mesh_id < 0 || mesh_id >= (int)c->meshes.size() || !c->meshes[mesh_id].live)
    return set_error(c, HGX_ERR_BAD_HANDLE, "mesh_release: unknown mesh %d", mesh_id);
  if (int rc = set_device(c)) return rc;
  HGX_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  free_mesh(c->meshes[mesh_id]);
  return HGX_OK;
}

int hgx_mesh_info(const hgx_ctx* h, int mesh_id, int64_t* ntri, double* volume, double* radius) {
  const Context* c = ctx_of(h);
  if (!c) return HGX_ERR_BAD_HANDLE;
  if (mesh_id < 0 || mesh_id >= (int)c->meshes.size() || !c->meshes[mesh_id].live) return HGX_ERR_BAD_HANDLE;
  const Mesh& m = c->meshes[mesh_id];
  if (ntri) *ntri = m.ntri;
  if (volume) *volume = m.volume;
  if (radius) *radius = m.radius;
  return HGX_OK;
}

}  // extern "C"
