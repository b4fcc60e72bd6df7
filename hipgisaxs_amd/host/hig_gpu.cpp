// hig_gpu.cpp -- the reference's GPU-class interfaces on the hgx C ABI (see hig_gpu.hpp).
#include "hig_gpu.hpp"

#include <cmath>
#include <cstring>
#include <iostream>

namespace hig {

namespace {

// FNV-1a over the bytes of an array: detects a changed q-grid or mesh between calls.
unsigned long long hash_bytes(const void* p, size_t n, unsigned long long h = 1469598103934665603ULL) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ULL;
  return h;
}

}  // namespace

// ----- RotMatrix_t (include/numerics/matrix.hpp:35-90) ---------------------------------------------

RotMatrix_t::RotMatrix_t() : data_{1, 0, 0, 0, 1, 0, 0, 0, 1} {}

RotMatrix_t::RotMatrix_t(int axis, real_t a) {
  real_t c = std::cos(a), s = std::sin(a);
  const real_t x[9] = {1, 0, 0, 0, c, -s, 0, s, c};
  const real_t y[9] = {c, 0, -s, 0, 1, 0, s, 0, c};
  const real_t z[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
  const real_t* m = axis == 0 ? x : axis == 1 ? y : z;
  std::memcpy(data_, m, sizeof(data_));
}

RotMatrix_t RotMatrix_t::operator*(const RotMatrix_t& rhs) const {
  RotMatrix_t r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      real_t s = 0;
      for (int k = 0; k < 3; ++k) s += data_[3 * i + k] * rhs.data_[3 * k + j];
      r.data_[3 * i + j] = s;
    }
  return r;
}

vector3_t RotMatrix_t::operator*(const vector3_t& v) const {
  return vector3_t(data_[0] * v[0] + data_[1] * v[1] + data_[2] * v[2], data_[3] * v[0] + data_[4] * v[1] + data_[5] * v[2],
                   data_[6] * v[0] + data_[7] * v[1] + data_[8] * v[2]);
}

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

bool GpuContext::set_grid(unsigned nqy, const real_t* qx, const real_t* qy, unsigned nqz, const complex_t* qz,
                          const real_t* alpha_f, unsigned nrow) {
  if (!ctx_) return false;
  if (nqy == 0 || (nqz != nqy && nqz != 4 * nqy) || !qx || !qy || !qz) {
    std::cerr << "error: q-grid sizes must satisfy nqz = nqy or 4 nqy" << std::endl;
    return false;
  }
  unsigned long long h = hash_bytes(qx, nqy * sizeof(real_t));
  h = hash_bytes(qy, nqy * sizeof(real_t), h);
  h = hash_bytes(qz, nqz * sizeof(complex_t), h);
  if (alpha_f) h = hash_bytes(alpha_f, nrow * sizeof(real_t), h);
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

// ----- NumericFormFactorG ------------------------------------------------------------------------

NumericFormFactorG::NumericFormFactorG() {}
NumericFormFactorG::NumericFormFactorG(int) {}
NumericFormFactorG::NumericFormFactorG(int, int, int) {}
NumericFormFactorG::~NumericFormFactorG() {
  if (mesh_id_ >= 0 && gpu_.ok()) hgx_mesh_release(gpu_.get(), mesh_id_);
}

unsigned int NumericFormFactorG::run(int kind, const real_t* rec, int ntri, int method, complex_t*& ff, int nqy,
                                     real_t* qx, real_t* qy, int nqz, complex_t* qz, RotMatrix_t& rot,
                                     real_t& kernel_time) {
  ff = nullptr;
  if (!gpu_.ok() || ntri < 1 || nqy < 1 || nqz < 1) return 0;
  if (!gpu_.set_grid((unsigned)nqy, qx, qy, (unsigned)nqz, qz)) return 0;
  size_t width = kind == HGX_MESH_APPROX8 ? 8 : 9;
  unsigned long long h = hash_bytes(rec, (size_t)ntri * width * sizeof(real_t), (unsigned long long)kind + 7);
  if (h != mesh_hash_ || mesh_id_ < 0) {
    if (mesh_id_ >= 0) hgx_mesh_release(gpu_.get(), mesh_id_);
    mesh_id_ = -1;
    int rc = hgx_mesh_upload(gpu_.get(), kind, ntri, rec, &mesh_id_);
    if (rc != HGX_OK) return gpu_.fail("hgx_mesh_upload", rc), 0;
    mesh_hash_ = h;
  }
  int64_t n_used = 0;
  hgx_mesh_info(gpu_.get(), mesh_id_, &n_used, nullptr, nullptr);
  size_t bytes = (size_t)nqz * sizeof(complex_t);
  double* d_ff = static_cast<double*>(gpu_.scratch(bytes));
  if (!d_ff) return gpu_.fail("hgx_device_alloc", HGX_ERR_OOM), 0;
  gpu_.begin_timing();
  int rc = hgx_ff_numeric(gpu_.get(), mesh_id_, method, precision_, 0, rot.data(), nullptr, d_ff);
  gpu_.end_timing();
  if (rc != HGX_OK) return gpu_.fail("hgx_ff_numeric", rc), 0;
  ff = new (std::nothrow) complex_t[nqz];
  if (!ff) {
    std::cerr << "Memory allocation failed for ff. Requested size: " << nqz << std::endl;
    return 0;
  }
  rc = hgx_memcpy_d2h(gpu_.get(), ff, d_ff, bytes);
  if (rc != HGX_OK) {
    delete[] ff;
    ff = nullptr;
    return gpu_.fail("hgx_memcpy_d2h", rc), 0;
  }
  kernel_time = gpu_.last_kernel_ms();
  return (unsigned int)n_used;
}

unsigned int NumericFormFactorG::compute_approx_triangle(std::vector<real_t>& shape_def, complex_t*& ff, int nqy,
                                                         real_t* qx, real_t* qy, int nqz, complex_t* qz,
                                                         RotMatrix_t& rot, real_t& kernel_time) {
  int ntri = (int)(shape_def.size() / 8);  // T_PROP_SIZE_ (include/common/parameters.hpp:32)
  return run(HGX_MESH_APPROX8, shape_def.data(), ntri, HGX_FF_APPROX, ff, nqy, qx, qy, nqz, qz, rot, kernel_time);
}

unsigned int NumericFormFactorG::compute_exact_triangle(triangle_t* shape_def, int num_triangles, complex_t*& ff,
                                                        int nqy, real_t* qx, real_t* qy, int nqz, complex_t* qz,
                                                        RotMatrix_t& rot, real_t& kernel_time) {
  int saved = precision_;
  precision_ = HGX_PREC_FP64;
  unsigned int n = run(HGX_MESH_EXACT9, reinterpret_cast<const real_t*>(shape_def), num_triangles, HGX_FF_EXACT, ff,
                       nqy, qx, qy, nqz, qz, rot, kernel_time);
  precision_ = saved;
  return n;
}

// ----- AnalyticFormFactorG -----------------------------------------------------------------------

AnalyticFormFactorG::AnalyticFormFactorG() {}
AnalyticFormFactorG::~AnalyticFormFactorG() {}

bool AnalyticFormFactorG::init(unsigned int nqy, const real_t* qx, const real_t* qy, unsigned int nqz,
                               const complex_t* qz) {
  if (!gpu_.set_grid(nqy, qx, qy, nqz, qz)) return false;
  nqz_ = nqz;
  return true;
}

bool AnalyticFormFactorG::run(int shape, int ndim, const real_vec_t* vals[4], const real_vec_t* wts[4], real_t tau,
                              real_t eta, const RotMatrix_t& rot, const real_vec_t& transvec, complex_vec_t& ff) {
  if (!gpu_.ok() || nqz_ == 0) {
    std::cerr << "error: analytic form factor used before init()" << std::endl;
    return false;
  }
  hgx_shape_params p;
  std::memset(&p, 0, sizeof(p));
  std::vector<real_t> ones[4];
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
  p.flags = compat_;
  size_t bytes = (size_t)nqz_ * sizeof(complex_t);
  double* d_ff = static_cast<double*>(gpu_.scratch(bytes));
  if (!d_ff) return gpu_.fail("hgx_device_alloc", HGX_ERR_OOM);
  double tv[3] = {0, 0, 0};
  for (size_t i = 0; i < 3 && i < transvec.size(); ++i) tv[i] = transvec[i];
  int rc = hgx_ff_analytic(gpu_.get(), shape, &p, HGX_PREC_FP64, rot.data(), tv, d_ff);
  if (rc != HGX_OK) return gpu_.fail("hgx_ff_analytic", rc);
  ff.clear();
  ff.resize(nqz_);
  rc = hgx_memcpy_d2h(gpu_.get(), ff.data(), d_ff, bytes);
  if (rc != HGX_OK) return gpu_.fail("hgx_memcpy_d2h", rc);
  return true;
}

bool AnalyticFormFactorG::compute_box(const real_t tau, const real_t eta, const real_vec_t& x, const real_vec_t& dx,
                                      const real_vec_t& y, const real_vec_t& dy, const real_vec_t& z,
                                      const real_vec_t& dz, const RotMatrix_t& rot, const real_vec_t& t,
                                      complex_vec_t& ff) {
  const real_vec_t* v[4] = {&x, &y, &z, nullptr};
  const real_vec_t* w[4] = {&dx, &dy, &dz, nullptr};
  return run(HGX_SHAPE_BOX, 3, v, w, tau, eta, rot, t, ff);
}

bool AnalyticFormFactorG::compute_cube(const real_t tau, const real_t eta, const real_vec_t& x, const real_vec_t& dx,
                                       const RotMatrix_t& rot, const real_vec_t& t, complex_vec_t& ff) {
  const real_vec_t* v[4] = {&x, nullptr, nullptr, nullptr};
  const real_vec_t* w[4] = {&dx, nullptr, nullptr, nullptr};
  return run(HGX_SHAPE_CUBE, 1, v, w, tau, eta, rot, t, ff);
}

bool AnalyticFormFactorG::compute_cylinder(const real_t tau, const real_t eta, const real_vec_t& r,
                                           const real_vec_t& dr, const real_vec_t& h, const real_vec_t& dh,
                                           const RotMatrix_t& rot, const real_vec_t& t, complex_vec_t& ff) {
  const real_vec_t* v[4] = {&r, &h, nullptr, nullptr};
  const real_vec_t* w[4] = {&dr, &dh, nullptr, nullptr};
  return run(HGX_SHAPE_CYLINDER, 2, v, w, tau, eta, rot, t, ff);
}

bool AnalyticFormFactorG::compute_sphere(const real_vec_t& r, const real_vec_t& dr, const real_vec_t&,
                                         const real_vec_t&, const RotMatrix_t& rot, const real_vec_t& t,
                                         complex_vec_t& ff) {
  // the height parameter is parsed but not used by the reference's sphere (ff_ana_sphere.cpp:98-110)
  const real_vec_t* v[4] = {&r, nullptr, nullptr, nullptr};
  const real_vec_t* w[4] = {&dr, nullptr, nullptr, nullptr};
  return run(HGX_SHAPE_SPHERE, 1, v, w, 0.0, 0.0, rot, t, ff);
}

bool AnalyticFormFactorG::compute_prism3(const real_t tau, const real_t eta, const real_vec_t& l, const real_vec_t& dl,
                                         const real_vec_t& h, const real_vec_t& dh, const RotMatrix_t& rot,
                                         const real_vec_t& t, complex_vec_t& ff) {
  const real_vec_t* v[4] = {&l, &h, nullptr, nullptr};
  const real_vec_t* w[4] = {&dl, &dh, nullptr, nullptr};
  return run(HGX_SHAPE_PRISM3, 2, v, w, tau, eta, rot, t, ff);
}

bool AnalyticFormFactorG::compute_prism6(const real_t tau, const real_t eta, const real_vec_t& l, const real_vec_t& dl,
                                         const real_vec_t& h, const real_vec_t& dh, const RotMatrix_t& rot,
                                         const real_vec_t& t, complex_vec_t& ff) {
  const real_vec_t* v[4] = {&l, &h, nullptr, nullptr};
  const real_vec_t* w[4] = {&dl, &dh, nullptr, nullptr};
  return run(HGX_SHAPE_PRISM6, 2, v, w, tau, eta, rot, t, ff);
}

bool AnalyticFormFactorG::compute_pyramid(const real_t tau, const real_t eta, const real_vec_t& x,
                                          const real_vec_t& dx, const real_vec_t& y, const real_vec_t& dy,
                                          const real_vec_t& h, const real_vec_t& dh, const real_vec_t& b,
                                          const real_vec_t& db, const RotMatrix_t& rot, const real_vec_t& t,
                                          complex_vec_t& ff) {
  const real_vec_t* v[4] = {&x, &y, &h, &b};
  const real_vec_t* w[4] = {&dx, &dy, &dh, &db};
  return run(HGX_SHAPE_PYRAMID, 4, v, w, tau, eta, rot, t, ff);
}

// ----- StructureFactorG --------------------------------------------------------------------------

StructureFactorG::StructureFactorG() {}
StructureFactorG::~StructureFactorG() {}

bool StructureFactorG::init(unsigned int nqy, const real_t* qx, const real_t* qy, unsigned int nqz,
                            const complex_t* qz) {
  if (!gpu_.set_grid(nqy, qx, qy, nqz, qz)) return false;
  nqz_ = nqz;
  return true;
}

bool StructureFactorG::compute(std::string expt, vector3_t center, const vector3_t& a, const vector3_t& b,
                               const vector3_t& c, vector3_t repet, vector3_t scaling, RotMatrix_t rot,
                               real_t dw_sigma, int flags) {
  if (expt != "saxs" && expt != "gisaxs") return false;  // src/sf/sf.cpp:99-101
  if (!gpu_.ok() || nqz_ == 0) {
    std::cerr << "error: structure factor used before init()" << std::endl;
    return false;
  }
  double la[3], lb[3], lc[3], rep[3], ce[3];
  for (int k = 0; k < 3; ++k) {
    la[k] = a[k] * scaling[0];
    lb[k] = b[k] * scaling[1];
    lc[k] = c[k] * scaling[2];
    rep[k] = repet[k];
    ce[k] = center[k];
  }
  double* d_sf = static_cast<double*>(gpu_.scratch((size_t)nqz_ * sizeof(complex_t)));
  if (!d_sf) return gpu_.fail("hgx_device_alloc", HGX_ERR_OOM);
  int rc = hgx_sf_lattice(gpu_.get(), la, lb, lc, rep, ce, dw_sigma, flags, rot.data(), d_sf);
  if (rc != HGX_OK) return gpu_.fail("hgx_sf_lattice", rc);
  return true;
}

void StructureFactorG::get_sf(complex_t*& sf) {
  if (!sf || nqz_ == 0) return;
  void* d_sf = gpu_.scratch((size_t)nqz_ * sizeof(complex_t));
  int rc = hgx_memcpy_d2h(gpu_.get(), sf, d_sf, (size_t)nqz_ * sizeof(complex_t));
  if (rc != HGX_OK) gpu_.fail("hgx_memcpy_d2h", rc);
}

// ----- MultiLayerG -------------------------------------------------------------------------------

bool MultiLayerG::init(const std::vector<Layer>& layers) {
  if (layers.empty()) return false;
  layers_ = layers;
  return true;
}

bool MultiLayerG::propagation_coeffs(complex_vec_t& coeff, real_t k0, real_t alpha_i, int order,
                                     const real_vec_t& alpha_f, unsigned ncol) {
  if (order == -1) {  // src/model/multilayer.cpp:111-116
    std::cerr << "Error: shapes can't be buried inside the substrate." << std::endl;
    return false;
  }
  unsigned nrow = (unsigned)alpha_f.size();
  unsigned npix = nrow * ncol;
  if (!gpu_.ok() || npix == 0 || layers_.empty()) return false;
  real_vec_t zeros(npix, 0.0);
  complex_vec_t qz(npix, complex_t(0, 0));
  if (!gpu_.set_grid(npix, zeros.data(), zeros.data(), npix, qz.data(), alpha_f.data(), nrow)) return false;
  real_vec_t d, b, t;
  for (const Layer& L : layers_) {
    d.push_back(L.delta);
    b.push_back(L.beta);
    t.push_back(L.thickness);
  }
  size_t bytes = (size_t)4 * npix * sizeof(complex_t);
  double* d_fc = static_cast<double*>(gpu_.scratch(bytes));
  if (!d_fc) return gpu_.fail("hgx_device_alloc", HGX_ERR_OOM);
  int rc = hgx_fresnel(gpu_.get(), (int)layers_.size(), d.data(), b.data(), t.data(), order, k0, alpha_i, d_fc);
  if (rc != HGX_OK) return gpu_.fail("hgx_fresnel", rc);
  coeff.assign((size_t)4 * npix, complex_t(0, 0));
  rc = hgx_memcpy_d2h(gpu_.get(), coeff.data(), d_fc, bytes);
  if (rc != HGX_OK) return gpu_.fail("hgx_memcpy_d2h", rc);
  return true;
}

// ----- DWBA combination ----------------------------------------------------------------------------

bool dwba_intensity(GpuContext& gpu, unsigned npix, unsigned nq, const complex_t* fc, const complex_t* sf,
                    const complex_t* ff, real_t weight, real_t* I) {
  if (!gpu.ok() || !ff || !I || npix == 0 || (nq != npix && nq != 4 * npix)) return false;
  // the kernel reads the branch count from the context's grid: give it a grid of matching shape
  real_vec_t zeros(npix, 0.0);
  complex_vec_t qz(nq, complex_t(0, 0));
  if (!gpu.set_grid(npix, zeros.data(), zeros.data(), nq, qz.data())) return false;
  size_t cb = (size_t)nq * sizeof(complex_t);
  double* d_fc = fc ? static_cast<double*>(gpu.scratch(cb, 1)) : nullptr;
  double* d_sf = sf ? static_cast<double*>(gpu.scratch(cb, 2)) : nullptr;
  double* d_ff = static_cast<double*>(gpu.scratch(cb, 0));
  double* d_I = static_cast<double*>(gpu.scratch((size_t)npix * sizeof(real_t), 3));
  if (!d_ff || !d_I || (fc && !d_fc) || (sf && !d_sf)) return gpu.fail("hgx_device_alloc", HGX_ERR_OOM);
  int rc = HGX_OK;
  if (fc) rc |= hgx_memcpy_h2d(gpu.get(), d_fc, fc, cb);
  if (sf) rc |= hgx_memcpy_h2d(gpu.get(), d_sf, sf, cb);
  rc |= hgx_memcpy_h2d(gpu.get(), d_ff, ff, cb);
  rc |= hgx_memcpy_h2d(gpu.get(), d_I, I, (size_t)npix * sizeof(real_t));
  if (rc != HGX_OK) return gpu.fail("hgx_memcpy_h2d", rc);
  rc = hgx_dwba_intensity(gpu.get(), d_fc, d_sf, d_ff, weight, 1, d_I);
  if (rc != HGX_OK) return gpu.fail("hgx_dwba_intensity", rc);
  rc = hgx_memcpy_d2h(gpu.get(), I, d_I, (size_t)npix * sizeof(real_t));
  if (rc != HGX_OK) return gpu.fail("hgx_memcpy_d2h", rc);
  return true;
}

}  // namespace hig
