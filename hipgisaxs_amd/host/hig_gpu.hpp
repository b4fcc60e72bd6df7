// hig_gpu.hpp -- drop-in C++ classes with the reference's GPU-class interfaces, implemented on the
// hgx C ABI (include/hgx.h). A HipGISAXS build links libhgx_hig.so + libhgx.so in place of its CUDA
// objects; the callers (FormFactor::compute_form_factor, StructureFactor, run_gisaxs) keep their
// shape. Every method returns the reference's success/failure value (counts or bool), prints the
// reason to std::cerr like the reference does, and never exits the process.
#pragma once

#include <complex>
#include <string>
#include <vector>

#include "hgx.h"

namespace hig {

typedef double real_t;                  // DOUBLEP build (include/common/typedefs.hpp:47-65)
typedef std::complex<real_t> complex_t; // layout-identical to cuDoubleComplex
typedef std::vector<real_t> real_vec_t;
typedef std::vector<complex_t> complex_vec_t;

// include/common/typedefs.hpp:102-106
typedef struct Triangle {
  real_t v1[3];
  real_t v2[3];
  real_t v3[3];
} triangle_t;

class vector3_t {
 public:
  real_t vec_[3];
  vector3_t() : vec_{0, 0, 0} {}
  vector3_t(real_t a, real_t b, real_t c) : vec_{a, b, c} {}
  real_t& operator[](int i) { return vec_[i]; }
  real_t operator[](int i) const { return vec_[i]; }
};

// Row-major rotation, include/numerics/matrix.hpp:30-140: RotMatrix_t(axis, angle) and products.
class RotMatrix_t {
 public:
  RotMatrix_t();
  RotMatrix_t(int axis, real_t angle);
  RotMatrix_t operator*(const RotMatrix_t& rhs) const;
  vector3_t operator*(const vector3_t& rhs) const;
  const real_t* data() const { return data_; }
  real_t data_[9];
};

// One hgx context plus the q-grid cache shared by the adapters below.
class GpuContext {
 public:
  explicit GpuContext(int device = 0);
  ~GpuContext();
  GpuContext(const GpuContext&) = delete;
  GpuContext& operator=(const GpuContext&) = delete;
  hgx_ctx* get() const { return ctx_; }
  bool ok() const { return ctx_ != nullptr; }
  // Uploads (qx, qy [nqy], qz [nqz]) unless identical to the last upload; nqz must be nqy or 4*nqy.
  bool set_grid(unsigned nqy, const real_t* qx, const real_t* qy, unsigned nqz, const complex_t* qz,
                const real_t* alpha_f = nullptr, unsigned nrow = 1);
  // Device scratch of at least bytes (grows).
  void* scratch(size_t bytes, int slot = 0);
  bool fail(const char* where, int rc) const;
  double last_kernel_ms() const { return last_ms_; }
  void begin_timing();
  void end_timing();

 private:
  hgx_ctx* ctx_ = nullptr;
  unsigned long long grid_hash_ = 0;
  unsigned nqy_ = 0, nqz_ = 0;
  void* scratch_[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t scratch_bytes_[4] = {0, 0, 0, 0};
  double last_ms_ = 0.0;
};

// Replaces NumericFormFactorG (include/ff/gpu/ff_num_gpu.cuh:31-157). The block-size constructor
// arguments of the reference are accepted and ignored (the launch shape is chosen by hgx).
class NumericFormFactorG {
 public:
  NumericFormFactorG();
  explicit NumericFormFactorG(int block_cuda);
  NumericFormFactorG(int block_cuda_t, int block_cuda_y, int block_cuda_z);
  ~NumericFormFactorG();

  // ff_tri_gpu.cu:395-521: shape_def = 8 reals per triangle (s, nx, ny, nz, cx, cy, cz, pad);
  // allocates ff = new complex_t[nqz] (the caller delete[]s it); returns the triangle count, 0 on
  // error; kernel_time in ms.
  unsigned int compute_approx_triangle(std::vector<real_t>& shape_def, complex_t*& ff, int nqy, real_t* qx,
                                       real_t* qy, int nqz, complex_t* qz, RotMatrix_t& rot, real_t& kernel_time);
  // ff_tri_cpu.cpp:159-195 (exact method); same ownership and return convention.
  unsigned int compute_exact_triangle(triangle_t* shape_def, int num_triangles, complex_t*& ff, int nqy,
                                      real_t* qx, real_t* qy, int nqz, complex_t* qz, RotMatrix_t& rot,
                                      real_t& kernel_time);
  // Extension: HGX_PREC_FP64 (default, plain) or HGX_PREC_FP32C.
  void precision(int p) { precision_ = p; }

 private:
  unsigned int run(int kind, const real_t* rec, int ntri, int method, complex_t*& ff, int nqy, real_t* qx,
                   real_t* qy, int nqz, complex_t* qz, RotMatrix_t& rot, real_t& kernel_time);
  GpuContext gpu_;
  int precision_ = HGX_PREC_FP64;
  int mesh_id_ = -1;
  unsigned long long mesh_hash_ = 0;
};

// Replaces AnalyticFormFactorG (include/ff/gpu/ff_ana_gpu.cuh:28-96). init() takes the q-grid
// arrays explicitly (the reference reads the QGrid singleton, src/ff/gpu/ff_ana_gpu.cu:57-98).
// Weights are applied for every shape (the reference's CPU ignores them for cylinder and prisms:
// pass unit weights to reproduce); tau/eta are the prism tilt angles.
class AnalyticFormFactorG {
 public:
  AnalyticFormFactorG();
  ~AnalyticFormFactorG();
  bool init(unsigned int nqy, const real_t* qx, const real_t* qy, unsigned int nqz, const complex_t* qz);
  bool clear() { return true; }
  bool destroy() { return true; }

  bool compute_box(const real_t tau, const real_t eta, const real_vec_t& x, const real_vec_t& distr_x,
                   const real_vec_t& y, const real_vec_t& distr_y, const real_vec_t& z, const real_vec_t& distr_z,
                   const RotMatrix_t& rot, const real_vec_t& transvec, complex_vec_t& ff);
  bool compute_cube(const real_t tau, const real_t eta, const real_vec_t& x, const real_vec_t& distr_x,
                    const RotMatrix_t& rot, const real_vec_t& transvec, complex_vec_t& ff);
  bool compute_cylinder(const real_t tau, const real_t eta, const real_vec_t& r, const real_vec_t& distr_r,
                        const real_vec_t& h, const real_vec_t& distr_h, const RotMatrix_t& rot,
                        const real_vec_t& transvec, complex_vec_t& ff);
  bool compute_sphere(const real_vec_t& r, const real_vec_t& distr_r, const real_vec_t& h, const real_vec_t& distr_h,
                      const RotMatrix_t& rot, const real_vec_t& transvec, complex_vec_t& ff);
  bool compute_prism3(const real_t tau, const real_t eta, const real_vec_t& l, const real_vec_t& distr_l,
                      const real_vec_t& h, const real_vec_t& distr_h, const RotMatrix_t& rot,
                      const real_vec_t& transvec, complex_vec_t& ff);
  bool compute_prism6(const real_t tau, const real_t eta, const real_vec_t& l, const real_vec_t& distr_l,
                      const real_vec_t& h, const real_vec_t& distr_h, const RotMatrix_t& rot,
                      const real_vec_t& transvec, complex_vec_t& ff);
  bool compute_pyramid(const real_t tau, const real_t eta, const real_vec_t& x, const real_vec_t& distr_x,
                       const real_vec_t& y, const real_vec_t& distr_y, const real_vec_t& h, const real_vec_t& distr_h,
                       const real_vec_t& b, const real_vec_t& distr_b, const RotMatrix_t& rot,
                       const real_vec_t& transvec, complex_vec_t& ff);
  // Extension: HGX_SHAPE_COMPAT_* bits (e.g. the reference's pyramid expression).
  void compat(int flags) { compat_ = flags; }

 private:
  bool run(int shape, int ndim, const real_vec_t* vals[4], const real_vec_t* wts[4], real_t tau, real_t eta,
           const RotMatrix_t& rot, const real_vec_t& transvec, complex_vec_t& ff);
  GpuContext gpu_;
  unsigned nqz_ = 0;
  int compat_ = 0;
};

// Replaces StructureFactorG (include/sf/gpu/sf_gpu.cuh:30-70). The reference's compute() takes a
// Lattice*; here the three lattice vectors a, b, c (Lattice::construct_vectors,
// src/model/structure.cpp:86-182) are passed, then scaled by `scaling` (src/sf/sf.cpp:169-171).
class StructureFactorG {
 public:
  StructureFactorG();
  ~StructureFactorG();
  bool init(unsigned int nqy, const real_t* qx, const real_t* qy, unsigned int nqz, const complex_t* qz);
  bool compute(std::string expt, vector3_t center, const vector3_t& a, const vector3_t& b, const vector3_t& c,
               vector3_t repet, vector3_t scaling, RotMatrix_t rot, real_t dw_sigma = 0.0, int flags = 0);
  // Copies the last result into sf (caller-allocated, nqz elements), as sf.cpp:297-305 expects.
  void get_sf(complex_t*& sf);

 private:
  GpuContext gpu_;
  unsigned nqz_ = 0;
};

// Replaces MultiLayer::propagation_coeffs (src/model/multilayer.cpp:110-155) on explicit rows.
class MultiLayerG {
 public:
  struct Layer {
    real_t delta, beta, thickness;
  };
  // layers below the vacuum, top to bottom, substrate last (MultiLayer::init order)
  bool init(const std::vector<Layer>& layers);
  // fc (4 * nrow * ncol) for the rows' alpha_f; order = the particle's layer (0 = vacuum).
  bool propagation_coeffs(complex_vec_t& coeff, real_t k0, real_t alpha_i, int order, const real_vec_t& alpha_f,
                          unsigned ncol);

 private:
  GpuContext gpu_;
  std::vector<Layer> layers_;
};

// The DWBA sum of run_gisaxs (src/sim/hipgisaxs_main.cpp:1084-1100) on host arrays:
// I[p] += |weight * sum_b fc_b sf_b ff_b|^2 (nbranch = nq / npix).
bool dwba_intensity(GpuContext& gpu, unsigned npix, unsigned nq, const complex_t* fc, const complex_t* sf,
                    const complex_t* ff, real_t weight, real_t* I);

}  // namespace hig
