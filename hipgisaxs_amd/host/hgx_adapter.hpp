// hgx_adapter.hpp -- the host side behind the reference's GPU classes, on plain types only.
//
// libhgx_adapter.so implements, over the hgx C ABI (include/hgx.h), what NumericFormFactorG,
// AnalyticFormFactorG and StructureFactorG do in the reference (include/ff/gpu/ff_num_gpu.cuh,
// include/ff/gpu/ff_ana_gpu.cuh, include/sf/gpu/sf_gpu.cuh). Its interface uses doubles,
// std::complex<double> and std::vector only, so it has no dependence on the reference's own type
// definitions: the classes with the reference's names and signatures are header-only wrappers in
// hig_gpu.hpp, compiled inside the reference's translation units against the reference's headers.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "hgx.h"

namespace hgx_adapter {

typedef std::complex<double> cplx;

// One hgx context plus the q-grid last uploaded to it.
class GpuContext {
 public:
  explicit GpuContext(int device = 0);
  ~GpuContext();
  GpuContext(const GpuContext&) = delete;
  GpuContext& operator=(const GpuContext&) = delete;
  hgx_ctx* get() const { return ctx_; }
  bool ok() const { return ctx_ != nullptr; }
  // Uploads (qx, qy [nqy], qz [nqz]) unless identical to the last upload; nqz must be nqy or 4*nqy.
  bool set_grid(unsigned nqy, const double* qx, const double* qy, unsigned nqz, const cplx* qz,
                const double* alpha_f = nullptr, unsigned nrow = 1);
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

// Numeric form factor of a triangle mesh (compute_approx_triangle / compute_exact_triangle): the
// mesh is uploaded once and reused while its bytes are unchanged.
class NumericFF {
 public:
  NumericFF() = default;
  ~NumericFF();
  // kind HGX_MESH_APPROX8 (8 doubles per triangle) or HGX_MESH_EXACT9 (9), method HGX_FF_APPROX or
  // HGX_FF_EXACT; writes nqz complex values (interleaved doubles) to ff_out; returns ntri on success,
  // 0 on error (message on std::cerr); kernel_ms: the kernels' time.
  unsigned run(int kind, const double* rec, int ntri, int method, int precision, double* ff_out, int nqy,
               const double* qx, const double* qy, int nqz, const double* qz, const double rot[9],
               double* kernel_ms);

 private:
  GpuContext gpu_;
  int mesh_id_ = -1;
  unsigned long long mesh_hash_ = 0;
};

// Analytic form factors of the reference's shapes on a grid given once (init).
class AnalyticFF {
 public:
  bool init(unsigned nqy, const double* qx, const double* qy, unsigned nqz, const cplx* qz);
  void clear() { nqz_ = 0; }
  unsigned nqz() const { return nqz_; }
  // ndim parameter dimensions of shape (HGX_SHAPE_*): values and weights (a single weight for one
  // value means stat none); writes nqz values into ff (resized).
  bool run(int shape, int ndim, const std::vector<double>* vals[4], const std::vector<double>* wts[4], double tau,
           double eta, int flags, const double rot[9], const std::vector<double>& transvec, std::vector<cplx>& ff);

 private:
  GpuContext gpu_;
  unsigned nqz_ = 0;
};

// Lattice structure factor (StructureFactorG): init with the grid, compute, then copy out.
class StructureF {
 public:
  bool init(unsigned nqy, const double* qx, const double* qy, unsigned nqz, const cplx* qz);
  bool ready() const { return nqz_ != 0; }
  void destroy() { nqz_ = 0; }
  // la, lb, lc already scaled; center: the grain centre; rot row-major
  bool compute(const std::string& expt, const double center[3], const double la[3], const double lb[3],
               const double lc[3], const double repet[3], const double rot[9], double dw_sigma, int flags);
  // copies the last result (nqz values) into sf
  bool get_sf(cplx* sf);

 private:
  GpuContext gpu_;
  unsigned nqz_ = 0;
};

// MultiLayer::propagation_coeffs (src/model/multilayer.cpp:110-155) on explicit rows.
class MultiLayer {
 public:
  struct Layer {
    double delta, beta, thickness;
  };
  bool init(const std::vector<Layer>& layers);
  // fc (4 * nrow * ncol) for the rows' alpha_f; order = the particle's layer (0 = vacuum).
  bool propagation_coeffs(std::vector<cplx>& coeff, double k0, double alpha_i, int order,
                          const std::vector<double>& alpha_f, unsigned ncol);

 private:
  GpuContext gpu_;
  std::vector<Layer> layers_;
};

// The DWBA sum of run_gisaxs (src/sim/hipgisaxs_main.cpp:1084-1100) on host arrays:
// I[p] += |weight * sum_b fc_b sf_b ff_b|^2 (nbranch = nq / npix).
bool dwba_intensity(GpuContext& gpu, unsigned npix, unsigned nq, const cplx* fc, const cplx* sf, const cplx* ff,
                    double weight, double* I);

}  // namespace hgx_adapter
