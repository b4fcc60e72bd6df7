// hig_gpu.hpp -- drop-in replacements of the reference's GPU classes NumericFormFactorG,
// AnalyticFormFactorG and StructureFactorG, with the reference's names and signatures
// (include/ff/gpu/ff_num_gpu.cuh:31-157, include/ff/gpu/ff_ana_gpu.cuh:28-96,
// include/sf/gpu/sf_gpu.cuh:30-70), implemented on libhgx_adapter.so (hgx_adapter.hpp) and the hgx
// C ABI.
//
// This header is included inside the reference's own translation units, in place of the three CUDA
// headers. It defines none of the reference's types: real_t, complex_t, cucomplex_t, triangle_t,
// vector3_t, RotMatrix_t, Lattice and QGrid come from the reference's headers included below, and
// the classes are header-only, so they compile against exactly the definitions the calling code
// sees. What crosses into the library is plain doubles. The reference is built in double precision
// (DOUBLEP): real_t must be double and cucomplex_t two doubles.
//
// Ownership and error conventions are the reference's: compute_approx_triangle and
// compute_exact_triangle allocate ff with new[] for the caller to delete[] and return the triangle
// count (0 on error); the others return bool; errors go to std::cerr; nothing exits the process.
// AnalyticFormFactorG::init and StructureFactorG::init read the q-grid from the QGrid singleton, as
// the reference's GPU classes do (src/ff/gpu/ff_ana_gpu.cu:57-98, src/sf/gpu/sf_gpu.cu:45-116).
#pragma once

#include <complex>
#include <iostream>
#include <new>
#include <string>
#include <vector>

#include <common/globals.hpp>    // vector3_t (include/common/globals.hpp:68-135)
#include <common/typedefs.hpp>   // real_t, complex_t, cucomplex_t, triangle_t (include/common/typedefs.hpp:38-106)
#include <model/qgrid.hpp>       // QGrid singleton (include/model/qgrid.hpp)
#include <model/structure.hpp>   // Lattice (include/model/structure.hpp:86-147)
#include <numerics/matrix.hpp>   // RotMatrix_t (include/numerics/matrix.hpp:30-140)

#include "hgx_adapter.hpp"

namespace hig {

static_assert(sizeof(real_t) == sizeof(double), "the hgx adapters take the reference's DOUBLEP build");
static_assert(sizeof(cucomplex_t) == 2 * sizeof(double), "cucomplex_t must be two doubles");
static_assert(sizeof(complex_t) == 2 * sizeof(double), "complex_t must be two doubles");
static_assert(sizeof(triangle_t) == 9 * sizeof(double), "triangle_t must be nine doubles");

namespace hgx_detail {

// The nine entries of a RotMatrix_t through its public interface (its data_ is private): column j is
// the rotation of the unit vector e_j.
inline void rot_entries(RotMatrix_t r, double m[9]) {
  for (int j = 0; j < 3; ++j) {
    std::vector<complex_t> col = r.rotate(j == 0 ? 1.0 : 0.0, j == 1 ? 1.0 : 0.0, complex_t(j == 2 ? 1.0 : 0.0, 0.0));
    for (int i = 0; i < 3; ++i) m[3 * i + j] = col[i].real();
  }
}

// The q-grid of the QGrid singleton: qx, qy per pixel (nqy), qz extended (nqz).
inline void singleton_grid(unsigned nqy, unsigned nqz, std::vector<double>& qx, std::vector<double>& qy,
                           std::vector<std::complex<double>>& qz) {
  qx.resize(nqy);
  qy.resize(nqy);
  qz.resize(nqz);
  for (unsigned i = 0; i < nqy; ++i) {
    qx[i] = QGrid::instance().qx(i);
    qy[i] = QGrid::instance().qy(i);
  }
  for (unsigned i = 0; i < nqz; ++i) {
    const complex_t z = QGrid::instance().qz_extended(i);
    qz[i] = std::complex<double>(z.real(), z.imag());
  }
}

inline std::vector<double> to_vec(const std::vector<real_t>& v) { return std::vector<double>(v.begin(), v.end()); }

}  // namespace hgx_detail

// include/ff/gpu/ff_num_gpu.cuh:31-157. The block shapes of the constructors are accepted and
// ignored (hgx chooses its launch shapes). precision() (HGX_PREC_FP64 by default, or HGX_PREC_FP32C)
// is an extension.
class NumericFormFactorG {
 public:
  explicit NumericFormFactorG(int /*block_cuda*/) {}
  NumericFormFactorG(int /*block_cuda_t*/, int /*block_cuda_y*/, int /*block_cuda_z*/) {}
  NumericFormFactorG() {}
  ~NumericFormFactorG() {}

  // src/ff/gpu/ff_tri_gpu.cu:395-521; shape_def: 8 reals per triangle (s, nx, ny, nz, cx, cy, cz, pad)
  unsigned int compute_approx_triangle(std::vector<real_t>& shape_def, cucomplex_t*& ff, int nqy, real_t* qx,
                                       real_t* qy, int nqz, cucomplex_t* qz, RotMatrix_t& rot, real_t& kernel_time) {
    return run(HGX_MESH_APPROX8, shape_def.data(), (int)(shape_def.size() / 8), HGX_FF_APPROX, precision_, ff, nqy,
               qx, qy, nqz, qz, rot, kernel_time);
  }
  // src/ff/gpu/ff_tri_gpu.cu:319-390 (exact method, FP64)
  unsigned int compute_exact_triangle(triangle_t* shape_def, int num_triangles, cucomplex_t*& ff, int nqy,
                                      real_t* qx, real_t* qy, int nqz, cucomplex_t* qz, RotMatrix_t& rot,
                                      real_t& kernel_time) {
    return run(HGX_MESH_EXACT9, reinterpret_cast<const real_t*>(shape_def), num_triangles, HGX_FF_EXACT,
               HGX_PREC_FP64, ff, nqy, qx, qy, nqz, qz, rot, kernel_time);
  }

  void precision(int p) { precision_ = p; }

 private:
  unsigned int run(int kind, const real_t* rec, int ntri, int method, int prec, cucomplex_t*& ff, int nqy,
                   real_t* qx, real_t* qy, int nqz, cucomplex_t* qz, RotMatrix_t& rot, real_t& kernel_time) {
    ff = nullptr;
    if (ntri < 1 || nqz < 1) return 0;
    double m[9];
    hgx_detail::rot_entries(rot, m);
    ff = new (std::nothrow) cucomplex_t[nqz];
    if (!ff) {
      std::cerr << "Memory allocation failed for ff. Requested size: " << nqz << std::endl;
      return 0;
    }
    double ms = 0.0;
    unsigned n = core_.run(kind, reinterpret_cast<const double*>(rec), ntri, method, prec,
                           reinterpret_cast<double*>(ff), nqy, qx, qy, nqz, reinterpret_cast<const double*>(qz), m,
                           &ms);
    if (n == 0) {
      delete[] ff;
      ff = nullptr;
      return 0;
    }
    kernel_time = ms;
    return n;
  }
  hgx_adapter::NumericFF core_;
  int precision_ = HGX_PREC_FP64;
};

// include/ff/gpu/ff_ana_gpu.cuh:28-96. Weights are applied for every shape (the reference's CPU
// ignores them for cylinders and prisms: pass unit weights to reproduce that). compat() (bits
// HGX_SHAPE_COMPAT_*) is an extension.
class AnalyticFormFactorG {
 public:
  AnalyticFormFactorG(unsigned int nqy, unsigned int nqz) { init(nqy, nqz); }
  AnalyticFormFactorG() {}
  ~AnalyticFormFactorG() {}

  // src/ff/gpu/ff_ana_gpu.cu:57-98: the grid from the QGrid singleton (call site src/ff/ff_ana.cpp:43)
  bool init(unsigned int nqy, unsigned int nqz) {
    std::vector<double> qx, qy;
    std::vector<std::complex<double>> qz;
    hgx_detail::singleton_grid(nqy, nqz, qx, qy, qz);
    return core_.init(nqy, qx.data(), qy.data(), nqz, qz.data());
  }
  bool clear() {
    core_.clear();
    return true;
  }
  bool destroy() { return clear(); }
  void grid_size(unsigned int nqy, unsigned int nqz) { init(nqy, nqz); }

  bool compute_box(const real_t tau, const real_t eta, const std::vector<real_t>& x, const std::vector<real_t>& dx,
                   const std::vector<real_t>& y, const std::vector<real_t>& dy, const std::vector<real_t>& z,
                   const std::vector<real_t>& dz, const RotMatrix_t& rot, const std::vector<real_t>& t,
                   std::vector<complex_t>& ff) {
    const std::vector<real_t>* v[4] = {&x, &y, &z, nullptr};
    const std::vector<real_t>* w[4] = {&dx, &dy, &dz, nullptr};
    return run(HGX_SHAPE_BOX, 3, v, w, tau, eta, rot, t, ff);
  }
  bool compute_cube(const real_t tau, const real_t eta, const std::vector<real_t>& x, const std::vector<real_t>& dx,
                    const RotMatrix_t& rot, const std::vector<real_t>& t, std::vector<complex_t>& ff) {
    const std::vector<real_t>* v[4] = {&x, nullptr, nullptr, nullptr};
    const std::vector<real_t>* w[4] = {&dx, nullptr, nullptr, nullptr};
    return run(HGX_SHAPE_CUBE, 1, v, w, tau, eta, rot, t, ff);
  }
  bool compute_cylinder(const real_t tau, const real_t eta, const std::vector<real_t>& r,
                        const std::vector<real_t>& dr, const std::vector<real_t>& h, const std::vector<real_t>& dh,
                        const RotMatrix_t& rot, const std::vector<real_t>& t, std::vector<complex_t>& ff) {
    const std::vector<real_t>* v[4] = {&r, &h, nullptr, nullptr};
    const std::vector<real_t>* w[4] = {&dr, &dh, nullptr, nullptr};
    return run(HGX_SHAPE_CYLINDER, 2, v, w, tau, eta, rot, t, ff);
  }
  // the height parameter is parsed but not used by the reference's sphere (src/ff/ff_ana_sphere.cpp:98-110)
  bool compute_sphere(const std::vector<real_t>& r, const std::vector<real_t>& dr, const std::vector<real_t>& /*h*/,
                      const std::vector<real_t>& /*dh*/, const RotMatrix_t& rot, const std::vector<real_t>& t,
                      std::vector<complex_t>& ff) {
    const std::vector<real_t>* v[4] = {&r, nullptr, nullptr, nullptr};
    const std::vector<real_t>* w[4] = {&dr, nullptr, nullptr, nullptr};
    return run(HGX_SHAPE_SPHERE, 1, v, w, 0.0, 0.0, rot, t, ff);
  }
  bool compute_prism3(const real_t tau, const real_t eta, const std::vector<real_t>& l, const std::vector<real_t>& dl,
                      const std::vector<real_t>& h, const std::vector<real_t>& dh, const RotMatrix_t& rot,
                      const std::vector<real_t>& t, std::vector<complex_t>& ff) {
    const std::vector<real_t>* v[4] = {&l, &h, nullptr, nullptr};
    const std::vector<real_t>* w[4] = {&dl, &dh, nullptr, nullptr};
    return run(HGX_SHAPE_PRISM3, 2, v, w, tau, eta, rot, t, ff);
  }
  bool compute_prism6(const real_t tau, const real_t eta, const std::vector<real_t>& l, const std::vector<real_t>& dl,
                      const std::vector<real_t>& h, const std::vector<real_t>& dh, const RotMatrix_t& rot,
                      const std::vector<real_t>& t, std::vector<complex_t>& ff) {
    const std::vector<real_t>* v[4] = {&l, &h, nullptr, nullptr};
    const std::vector<real_t>* w[4] = {&dl, &dh, nullptr, nullptr};
    return run(HGX_SHAPE_PRISM6, 2, v, w, tau, eta, rot, t, ff);
  }
  bool compute_pyramid(const real_t tau, const real_t eta, const std::vector<real_t>& x,
                       const std::vector<real_t>& dx, const std::vector<real_t>& y, const std::vector<real_t>& dy,
                       const std::vector<real_t>& h, const std::vector<real_t>& dh, const std::vector<real_t>& b,
                       const std::vector<real_t>& db, const RotMatrix_t& rot, const std::vector<real_t>& t,
                       std::vector<complex_t>& ff) {
    const std::vector<real_t>* v[4] = {&x, &y, &h, &b};
    const std::vector<real_t>* w[4] = {&dx, &dy, &dh, &db};
    return run(HGX_SHAPE_PYRAMID, 4, v, w, tau, eta, rot, t, ff);
  }

  void compat(int flags) { compat_ = flags; }

 private:
  bool run(int shape, int ndim, const std::vector<real_t>* vals[4], const std::vector<real_t>* wts[4], real_t tau,
           real_t eta, const RotMatrix_t& rot, const std::vector<real_t>& transvec, std::vector<complex_t>& ff) {
    std::vector<double> v[4], w[4];
    const std::vector<double>* pv[4] = {nullptr, nullptr, nullptr, nullptr};
    const std::vector<double>* pw[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int d = 0; d < ndim; ++d) {
      v[d] = hgx_detail::to_vec(*vals[d]);
      w[d] = hgx_detail::to_vec(*wts[d]);
      pv[d] = &v[d];
      pw[d] = &w[d];
    }
    double m[9];
    hgx_detail::rot_entries(rot, m);
    std::vector<std::complex<double>> out;
    if (!core_.run(shape, ndim, pv, pw, tau, eta, compat_, m, hgx_detail::to_vec(transvec), out)) return false;
    ff.assign(out.size(), complex_t(0, 0));
    for (size_t i = 0; i < out.size(); ++i) ff[i] = complex_t(out[i].real(), out[i].imag());
    return true;
  }
  hgx_adapter::AnalyticFF core_;
  int compat_ = 0;
};

// include/sf/gpu/sf_gpu.cuh:30-70. compute() takes the Lattice the reference passes and scales its
// vectors by `scaling` (src/sf/gpu/sf_gpu.cu:205-207). The lattice sum is the reference's CPU formula
// (src/sf/sf.cpp:165-245, no basis term 1 + exp(i q.t); every built-in lattice but fco has t = 0).
// dw_sigma() (a Debye-Waller factor) and flags() (HGX_SF_*) are extensions.
class StructureFactorG {
 public:
  StructureFactorG(unsigned int nx, unsigned int ny, unsigned int nz) { init(nx, ny, nz); }
  StructureFactorG() {}
  ~StructureFactorG() { destroy(); }

  // src/sf/gpu/sf_gpu.cu:45-116: the grid from the QGrid singleton (call sites src/sf/sf.cpp:58, 300)
  bool init(unsigned int /*nx*/, unsigned int ny, unsigned int nz) {
    if (core_.ready()) {
      std::cerr << "warning: gpu sf already initialized" << std::endl;
      return false;
    }
    std::vector<double> qx, qy;
    std::vector<std::complex<double>> qz;
    hgx_detail::singleton_grid(ny, nz, qx, qy, qz);
    nqz_ = nz;
    return core_.init(ny, qx.data(), qy.data(), nz, qz.data());
  }
  bool clear() { return destroy(); }
  bool destroy() {
    core_.destroy();
    nqz_ = 0;
    return true;
  }
  void grid_size(unsigned int nx, unsigned int ny, unsigned int nz) {
    destroy();
    init(nx, ny, nz);
  }

  // src/sf/gpu/sf_gpu.cu:182-241 (call site src/sf/sf.cpp:301-306)
  bool compute(std::string expt, vector3_t center, Lattice* lattice, vector3_t repet, vector3_t scaling,
               RotMatrix_t rot) {
    if (!lattice) return false;
    vector3_t a = lattice->a(), b = lattice->b(), c = lattice->c();
    double la[3], lb[3], lc[3], rep[3], ce[3], m[9];
    for (int k = 0; k < 3; ++k) {
      la[k] = a[k] * scaling[0];
      lb[k] = b[k] * scaling[1];
      lc[k] = c[k] * scaling[2];
      rep[k] = repet[k];
      ce[k] = center[k];
    }
    hgx_detail::rot_entries(rot, m);
    return core_.compute(expt, ce, la, lb, lc, rep, m, dw_sigma_, flags_);
  }
  // src/sf/gpu/sf_gpu.cu:118-140: copies the last result into the caller's nqz-element buffer
  void get_sf(complex_t*& sf) {
    if (!sf || nqz_ == 0) return;
    std::vector<std::complex<double>> out(nqz_);
    if (!core_.get_sf(out.data())) return;
    for (unsigned i = 0; i < nqz_; ++i) sf[i] = complex_t(out[i].real(), out[i].imag());
  }

  void dw_sigma(real_t s) { dw_sigma_ = s; }
  void flags(int f) { flags_ = f; }

 private:
  hgx_adapter::StructureF core_;
  unsigned nqz_ = 0;
  double dw_sigma_ = 0.0;
  int flags_ = 0;
};

}  // namespace hig
