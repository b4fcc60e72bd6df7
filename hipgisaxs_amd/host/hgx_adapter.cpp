// hgx_adapter.cpp -- plain-type host side of the reference's GPU classes on the hgx C ABI (see
// hgx_adapter.hpp; the reference-named classes are the header-only wrappers of hig_gpu.hpp).
#include "hgx_adapter.hpp"

#include <cmath>
#include <cstring>
#include <iostream>

namespace hgx_adapter {

namespace {

// FNV-1a over the bytes of an array: detects a changed q-grid or mesh between calls.
unsigned long long hash_bytes(const void* p, size_t n, unsigned long long h = 1469598103934665603ULL) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ULL;
  return h;
}

}  // namespace

// ----- GpuContext --------------------------------------------------------------------------------

GpuContext::GpuContext(int device) {
  int rc = hgx_create(device, &ctx_);
  if (rc != HGX_OK) {
    std::cerr << "error: hgx_create failed (" << rc << ")" << std::endl;
    ctx_ = nullptr;
  }
}

GpuContext::~GpuContext() {
  if (!ctx_) return;
  for (int i = 0; i < 4; ++i)
    if (scratch_[i]) hgx_device_free(ctx_, scratch_[i]);
  hgx_destroy(ctx_);
}

bool GpuContext::fail(const char* where, int rc) const {
  std::cerr << "error: " << where << " failed (" << rc << "): " << (ctx_ ? hgx_last_error(ctx_) : "no context")
            << std::endl;
  return false;
}

bool GpuContext::set_grid(unsigned nqy, const double* qx, const double* qy, unsigned nqz, const cplx* qz,
                          const double* alpha_f, unsigned nrow) {
  if (!ctx_) return false;
  if (nqy == 0 || (nqz != nqy && nqz != 4 * nqy) || !qx || !qy || !qz) {
    std::cerr << "error: q-grid sizes must satisfy nqz = nqy or 4 nqy" << std::endl;
    return false;
  }
  unsigned long long h = hash_bytes(qx, nqy * sizeof(double));
  h = hash_bytes(qy, nqy * sizeof(double), h);
  h = hash_bytes(qz, nqz * sizeof(cplx), h);
  if (alpha_f) h = hash_bytes(alpha_f, nrow * sizeof(double), h);
  if (h == grid_hash_ && nqy == nqy_ && nqz == nqz_) return true;
  unsigned ncol = nqy / nrow;
  if (ncol * nrow != nqy) {
    std::cerr << "error: nqy is not a multiple of nrow" << std::endl;
    return false;
  }
  int rc = hgx_qgrid_set(ctx_, (int)ncol, (int)nrow, (int)(nqz / nqy), qx, qy, reinterpret_cast<const double*>(qz),
                         alpha_f);
  if (rc != HGX_OK) return fail("hgx_qgrid_set", rc);
  grid_hash_ = h;
  nqy_ = nqy;
  nqz_ = nqz;
  return true;
}

void* GpuContext::scratch(size_t bytes, int slot) {
  if (bytes <= scratch_bytes_[slot]) return scratch_[slot];
  if (scratch_[slot]) hgx_device_free(ctx_, scratch_[slot]);
  scratch_[slot] = nullptr;
  scratch_bytes_[slot] = 0;
  if (hgx_device_alloc(ctx_, bytes, &scratch_[slot]) != HGX_OK) return nullptr;
  scratch_bytes_[slot] = bytes;
  return scratch_[slot];
}

void GpuContext::begin_timing() { hgx_profile_begin(ctx_); }

void GpuContext::end_timing() {
  double main_ms = 0, aux_ms = 0;
  int64_t nm = 0, na = 0;
  hgx_profile_end(ctx_, &main_ms, &nm, &aux_ms, &na);
  last_ms_ = main_ms + aux_ms;
}

// ----- NumericFF ------------------------------------------------------------------------------------

NumericFF::~NumericFF() {
  if (mesh_id_ >= 0 && gpu_.ok()) hgx_mesh_release(gpu_.get(), mesh_id_);
}

unsigned NumericFF::run(int kind, const double* rec, int ntri, int method, int precision, double* ff_out, int nqy,
                        const double* qx, const double* qy, int nqz, const double* qz, const double rot[9],
                        double* kernel_ms) {
  if (!gpu_.ok() || ntri < 1 || nqy < 1 || nqz < 1 || !rec || !ff_out) return 0;
  if (!gpu_.set_grid((unsigned)nqy, qx, qy, (unsigned)nqz, reinterpret_cast<const cplx*>(qz))) return 0;
  const size_t width = kind == HGX_MESH_APPROX8 ? 8 : 9;
  const unsigned long long h = hash_bytes(rec, (size_t)ntri * width * sizeof(double), (unsigned long long)kind + 7);
  if (h != mesh_hash_ || mesh_id_ < 0) {
    if (mesh_id_ >= 0) hgx_mesh_release(gpu_.get(), mesh_id_);
    mesh_id_ = -1;
    int rc = hgx_mesh_upload(gpu_.get(), kind, ntri, rec, &mesh_id_);
    if (rc != HGX_OK) return gpu_.fail("hgx_mesh_upload", rc), 0;
    mesh_hash_ = h;
  }
  const size_t bytes = (size_t)nqz * sizeof(cplx);
  double* d_ff = static_cast<double*>(gpu_.scratch(bytes));
  if (!d_ff) return gpu_.fail("hgx_device_alloc", HGX_ERR_OOM), 0;
  gpu_.begin_timing();
  int rc = hgx_ff_numeric(gpu_.get(), mesh_id_, method, precision, 0, rot, nullptr, d_ff);
  gpu_.end_timing();
  if (rc != HGX_OK) return gpu_.fail("hgx_ff_numeric", rc), 0;
  rc = hgx_memcpy_d2h(gpu_.get(), ff_out, d_ff, bytes);
  if (rc != HGX_OK) return gpu_.fail("hgx_memcpy_d2h", rc), 0;
  if (kernel_ms) *kernel_ms = gpu_.last_kernel_ms();
  return (unsigned)ntri;
}

// ----- AnalyticFF -----------------------------------------------------------------------------------

bool AnalyticFF::init(unsigned nqy, const double* qx, const double* qy, unsigned nqz, const cplx* qz) {
  if (!gpu_.set_grid(nqy, qx, qy, nqz, qz)) return false;
  nqz_ = nqz;
  return true;
}

bool AnalyticFF::run(int shape, int ndim, const std::vector<double>* vals[4], const std::vector<double>* wts[4],
                     double tau, double eta, int flags, const double rot[9], const std::vector<double>& transvec,
                     std::vector<cplx>& ff) {
  if (!gpu_.ok() || nqz_ == 0) {
    std::cerr << "error: analytic form factor used before init()" << std::endl;
    return false;
  }
  hgx_shape_params p;
  std::memset(&p, 0, sizeof(p));
  std::vector<double> ones[4];
  for (int d = 0; d < ndim; ++d) {
    if (vals[d]->empty()) {
      std::cerr << "error: missing parameter values for shape dimension " << d << std::endl;
      return false;
    }
    p.n[d] = (int)vals[d]->size();
    p.val[d] = vals[d]->data();
    if (wts[d]->size() >= vals[d]->size()) {
      p.wt[d] = wts[d]->data();
    } else {  // stat_none: one weight for a single value (src/ff/ff_ana.cpp:262-263)
      ones[d].assign(vals[d]->size(), 1.0);
      p.wt[d] = ones[d].data();
    }
  }
  p.tau = tau;
  p.eta = eta;
  p.flags = flags;
  const size_t bytes = (size_t)nqz_ * sizeof(cplx);
  double* d_ff = static_cast<double*>(gpu_.scratch(bytes));
  if (!d_ff) return gpu_.fail("hgx_device_alloc", HGX_ERR_OOM);
  double tv[3] = {0, 0, 0};
  for (size_t i = 0; i < 3 && i < transvec.size(); ++i) tv[i] = transvec[i];
  int rc = hgx_ff_analytic(gpu_.get(), shape, &p, HGX_PREC_FP64, rot, tv, d_ff);
  if (rc != HGX_OK) return gpu_.fail("hgx_ff_analytic", rc);
  ff.clear();
  ff.resize(nqz_);
  rc = hgx_memcpy_d2h(gpu_.get(), ff.data(), d_ff, bytes);
  if (rc != HGX_OK) return gpu_.fail("hgx_memcpy_d2h", rc);
  return true;
}

// ----- StructureF -----------------------------------------------------------------------------------

bool StructureF::init(unsigned nqy, const double* qx, const double* qy, unsigned nqz, const cplx* qz) {
  if (!gpu_.set_grid(nqy, qx, qy, nqz, qz)) return false;
  nqz_ = nqz;
  return true;
}

bool StructureF::compute(const std::string& expt, const double center[3], const double la[3], const double lb[3],
                         const double lc[3], const double repet[3], const double rot[9], double dw_sigma, int flags) {
  if (expt != "saxs" && expt != "gisaxs") return false;  // src/sf/sf.cpp:99-101
  if (!gpu_.ok() || nqz_ == 0) {
    std::cerr << "error: structure factor used before init()" << std::endl;
    return false;
  }
  double* d_sf = static_cast<double*>(gpu_.scratch((size_t)nqz_ * sizeof(cplx)));
  if (!d_sf) return gpu_.fail("hgx_device_alloc", HGX_ERR_OOM);
  int rc = hgx_sf_lattice(gpu_.get(), la, lb, lc, repet, center, dw_sigma, flags, rot, d_sf);
  if (rc != HGX_OK) return gpu_.fail("hgx_sf_lattice", rc);
  return true;
}

bool StructureF::get_sf(cplx* sf) {
  if (!sf || nqz_ == 0) return false;
  void* d_sf = gpu_.scratch((size_t)nqz_ * sizeof(cplx));
  int rc = hgx_memcpy_d2h(gpu_.get(), sf, d_sf, (size_t)nqz_ * sizeof(cplx));
  if (rc != HGX_OK) return gpu_.fail("hgx_memcpy_d2h", rc);
  return true;
}

// ----- MultiLayer -----------------------------------------------------------------------------------

bool MultiLayer::init(const std::vector<Layer>& layers) {
  if (layers.empty()) return false;
  layers_ = layers;
  return true;
}

bool MultiLayer::propagation_coeffs(std::vector<cplx>& coeff, double k0, double alpha_i, int order,
                                    const std::vector<double>& alpha_f, unsigned ncol) {
  if (order == -1) {  // src/model/multilayer.cpp:111-116
    std::cerr << "Error: shapes can't be buried inside the substrate." << std::endl;
    return false;
  }
  unsigned nrow = (unsigned)alpha_f.size();
  unsigned npix = nrow * ncol;
  if (!gpu_.ok() || npix == 0 || layers_.empty()) return false;
  std::vector<double> zeros(npix, 0.0);
  std::vector<cplx> qz(npix, cplx(0, 0));
  if (!gpu_.set_grid(npix, zeros.data(), zeros.data(), npix, qz.data(), alpha_f.data(), nrow)) return false;
  std::vector<double> d, b, t;
  for (const Layer& L : layers_) {
    d.push_back(L.delta);
    b.push_back(L.beta);
    t.push_back(L.thickness);
  }
  size_t bytes = (size_t)4 * npix * sizeof(cplx);
  double* d_fc = static_cast<double*>(gpu_.scratch(bytes));
  if (!d_fc) return gpu_.fail("hgx_device_alloc", HGX_ERR_OOM);
  int rc = hgx_fresnel(gpu_.get(), (int)layers_.size(), d.data(), b.data(), t.data(), order, k0, alpha_i, d_fc);
  if (rc != HGX_OK) return gpu_.fail("hgx_fresnel", rc);
  coeff.assign((size_t)4 * npix, cplx(0, 0));
  rc = hgx_memcpy_d2h(gpu_.get(), coeff.data(), d_fc, bytes);
  if (rc != HGX_OK) return gpu_.fail("hgx_memcpy_d2h", rc);
  return true;
}

// ----- DWBA combination ----------------------------------------------------------------------------

bool dwba_intensity(GpuContext& gpu, unsigned npix, unsigned nq, const cplx* fc, const cplx* sf,
                    const cplx* ff, double weight, double* I) {
  if (!gpu.ok() || !ff || !I || npix == 0 || (nq != npix && nq != 4 * npix)) return false;
  // the kernel reads the branch count from the context's grid: give it a grid of matching shape
  std::vector<double> zeros(npix, 0.0);
  std::vector<cplx> qz(nq, cplx(0, 0));
  if (!gpu.set_grid(npix, zeros.data(), zeros.data(), nq, qz.data())) return false;
  size_t cb = (size_t)nq * sizeof(cplx);
  double* d_fc = fc ? static_cast<double*>(gpu.scratch(cb, 1)) : nullptr;
  double* d_sf = sf ? static_cast<double*>(gpu.scratch(cb, 2)) : nullptr;
  double* d_ff = static_cast<double*>(gpu.scratch(cb, 0));
  double* d_I = static_cast<double*>(gpu.scratch((size_t)npix * sizeof(double), 3));
  if (!d_ff || !d_I || (fc && !d_fc) || (sf && !d_sf)) return gpu.fail("hgx_device_alloc", HGX_ERR_OOM);
  int rc = HGX_OK;
  if (fc) rc |= hgx_memcpy_h2d(gpu.get(), d_fc, fc, cb);
  if (sf) rc |= hgx_memcpy_h2d(gpu.get(), d_sf, sf, cb);
  rc |= hgx_memcpy_h2d(gpu.get(), d_ff, ff, cb);
  rc |= hgx_memcpy_h2d(gpu.get(), d_I, I, (size_t)npix * sizeof(double));
  if (rc != HGX_OK) return gpu.fail("hgx_memcpy_h2d", rc);
  rc = hgx_dwba_intensity(gpu.get(), d_fc, d_sf, d_ff, weight, 1, d_I);
  if (rc != HGX_OK) return gpu.fail("hgx_dwba_intensity", rc);
  rc = hgx_memcpy_d2h(gpu.get(), I, d_I, (size_t)npix * sizeof(double));
  if (rc != HGX_OK) return gpu.fail("hgx_memcpy_d2h", rc);
  return true;
}

}  // namespace hgx_adapter
