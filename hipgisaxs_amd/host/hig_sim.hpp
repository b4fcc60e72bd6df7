// hig_sim.hpp -- the simulation driver of a .hig input on the hgx C ABI: HipGISAXS::init,
// compute_gisaxs, run_gisaxs and run_all_gisaxs (src/sim/hipgisaxs_main.cpp:106-216, 283-626,
// 630-1436) with the same loop structure (structures -> grains -> unit-cell elements ->
// locations), reference orientation sampling and output layout, every array on the device.
#pragma once

#include <array>
#include <cstdint>
#include <exception>
#include <new>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "hgx.h"
#include "hgx_hig.h"
#include "hgx_adapter.hpp"
#include "hig_input.hpp"

namespace hig {

// The driver's rotations (row-major), with the reference's constructions and products
// (include/numerics/matrix.hpp:35-110): about axis 0, 1 or 2 by an angle, matrix products summed in
// the reference's order, rotation of a vector.
struct Rot3 {
  double m[9];
  Rot3();
  Rot3(int axis, double angle);
  Rot3 operator*(const Rot3& rhs) const;
  std::array<double, 3> operator*(const std::array<double, 3>& v) const;
  const double* data() const { return m; }
};

// The random() generator of glibc (additive feedback, degree 31: r[i] = r[i-3] + r[i-31] mod 2^32
// after an LCG seeding of r[0..30] and 310 discarded outputs; rand() returns r >> 1). A process
// that never calls srand() draws from seed 1; the reference samples grain orientations with rand()
// (src/sim/hipgisaxs_main.cpp:1917-1936) and seeds nothing on that path.
class GlibcRandom {
 public:
  explicit GlibcRandom(unsigned seed = 1) { reseed(seed); }
  void reseed(unsigned seed);
  int next();  // in [0, 2^31 - 1]
  static constexpr double kRandMax = 2147483647.0;

 private:
  uint32_t r_[34];
  int i_ = 0;  // ring position of r[i - 34]
};

// The uniform draw of woo::MTRandomNumberGenerator (include/woo/random/woo_mtrandom.hpp:28-33):
// (mt19937_64() - min) / (max - min).
struct MtUniform {
  std::mt19937_64 g;
  explicit MtUniform(uint64_t seed) : g(seed) {}
  double next() { return (double)(g() - std::mt19937_64::min()) / (double)(std::mt19937_64::max() - std::mt19937_64::min()); }
};

struct HigGrain {
  Rot3 rot;             // rot_phi * r3 * r2 * r1
  double center[3];     // rot * dd + transvec
  double weight;        // gauss_weight * scaling weight
};

// One step of run_gisaxs's grain scaling / repetition loop (src/sim/hipgisaxs_main.cpp:1031-1061)
// over the grains it applies to: the lattice scaling and repetitions of the step, each grain's
// weight multiplied by the step's scaling weight.
struct GrainSet {
  vec3 scaling;
  vec3 repet;
  std::vector<HigGrain> grains;
};

// The hgx_ensemble_model of one structure (an analytic unit cell seen by its grains) with the host
// arrays it points into.
struct EnsembleBuild {
  std::vector<hgx_ensemble_element> el;
  std::vector<std::vector<double>> stores;  // 8 per element: node values and weights per dimension
  std::vector<std::vector<double>> locs;    // per element [nloc][3]
  std::vector<double> rots, cen, wts;       // per grain
  hgx_ensemble_model m{};
  EnsembleBuild() = default;
  EnsembleBuild(const EnsembleBuild&) = delete;
  EnsembleBuild& operator=(const EnsembleBuild&) = delete;
  void relink();  // points m and el at the arrays above
};

// Parsing, configurations and grain sampling run on the host alone; the device context is created
// by the first call that needs it (init, the runs).
class HipGISAXS {
 public:
  explicit HipGISAXS(int device = 0);
  ~HipGISAXS();
  HipGISAXS(const HipGISAXS&) = delete;
  HipGISAXS& operator=(const HipGISAXS&) = delete;

  bool construct_input(const std::string& path);
  bool construct_input_text(const std::string& text, const std::string& base_dir = ".");
  // precision HGX_PREC_FP64 / HGX_PREC_FP32C, HGX_HIG_* flags, orientation sampling seed (1 = a
  // fresh reference process)
  void options(int precision, int flags, unsigned seed);
  // photon energy -> k0 and the image size (host only)
  bool prepare();
  // prepare, then the q-grid of the output region at the smallest incidence angle on the device
  bool init();
  // One image [nrow][ncol] for incidence alpha_i and in-plane rotation phi (degrees).
  bool run_gisaxs(double alpha_i_deg, double phi_deg, double tilt_deg, std::vector<double>& img);
  // The single-configuration run (fails when the input scans alpha_i, phi or tilt).
  bool compute_gisaxs(std::vector<double>& img);
  // The (alpha_i, phi, tilt) scan of run_all_gisaxs, in its loop order.
  std::vector<std::array<double, 3>> configurations() const;
  // The grains the next run would use for structure `key` (does not advance the sampler).
  bool grains(const std::string& key, std::vector<HigGrain>& out);
  static bool save_gisaxs(const double* data, unsigned ncol, unsigned nrow, const std::string& path);
  // The image file of run_all_gisaxs (Image::construct_image + Image::save, src/image/image.cpp:125-226,
  // 502-512): pixels translated to non-negative, log10 (zero stays zero), normalised to [0, 1]; written
  // as an uncompressed 16-bit grayscale TIFF (the reference maps [0, 1] through its colour palette into
  // an 8-bit RGB TIFF). Non-finite pixels are an error.
  static bool save_tiff(const double* data, unsigned ncol, unsigned nrow, const std::string& path);

  // ---- fitting: the HipGISAXSObjectiveFunction of the reference's analyzer
  // (src/analyzer/objective_func_hipgisaxs.cpp:47-72, 110-170, 245-332) on the device.
  // Sets the model value a fit variable path names (HiGInput::update_params); a change of an
  // instrument value re-derives k0 and the q-grid at the next run.
  bool update_param(const std::string& variable, double value);
  // Reads the input's reference data (ASCII rows, ImageData::read) and mask (whitespace-separated
  // weights; none: all 1), overrides the output region with the reference region when the sizes
  // differ (override_qregion), and keeps both on the device.
  bool fit_init();
  bool fit_ready() const { return fit_ready_; }
  // The fit vector's keys (sorted, as the reference's maps order them), limits, steps and initial values.
  std::vector<std::string> fit_keys() const;
  // The objective of x [nkeys]: update every parameter, compute_gisaxs on the device, the distance
  // under `metric` (HGX_METRIC_*) to the reference data, plus (reg/2) |x - range midpoints|^2; a
  // non-finite distance becomes 1e8. Appends (x, value) to the evaluation history.
  bool fit_objective(const double* x, int metric, double reg, double* value);
  // The objective at x and its gradient in one pass: forward-mode derivatives through the ensemble
  // kernel (hgx_ensemble_intensity_tangent) for every variable that moves only the model's tables
  // (shape parameters and rotations, contrasts, locations, lattice vectors and scaling, grain
  // centres, orientation angles and weights, iratio), with the tables' tangents from host central
  // differences of their construction; the distance's derivative from hgx_image_distance_tangent.
  // computed[j] = 0 for a variable it cannot differentiate so (instrument and layer values, which
  // move the q-grid or the Fresnel coefficients; integer counts; cells with numeric shapes): the
  // caller differentiates those otherwise. The value is the FP64 objective. Appends (x, value) to
  // the history.
  bool fit_gradient(const double* x, int metric, double reg, double* value, double* grad, int* computed);
  // fit_objective of every row of X in turn, the images and distances queued back to back and the
  // distances read back with one synchronisation (the generation of the particle swarm,
  // src/analyzer/hipgisaxs_fit_pso.cpp:521-543); same values and history as the calls one by one.
  bool fit_objective_many(const std::vector<std::vector<double>>& X, int metric, double reg, std::vector<double>& out);
  std::vector<std::vector<double>>& fit_history() { return history_; }

  int flags() const { return flags_; }
  unsigned nrows() const { return nrow_; }
  unsigned ncols() const { return ncol_; }
  double k0() const { return k0_; }
  const std::string& error() const { return err_; }
  const HigInput& input() const { return in_; }
  // work counts of the last run: grains through the fused kernel, grains through per-grain
  // calls, form-factor and structure-factor evaluations
  void last_counts(int64_t counts[4]) const;
  // numeric form-factor calls of the last run by kernel path: per-pair, product grid, Ewald grid
  void last_paths(int64_t paths[3]) const {
    for (int i = 0; i < 3; ++i) paths[i] = paths_[i];
  }
  double last_run_ms() const { return last_ms_; }
  // kernel times of the last run (HGX_HIG_PROFILE): main ms, main launches, aux ms, aux launches
  void last_kernel_times(double* main_ms, int64_t* main_n, double* aux_ms, int64_t* aux_n) const {
    *main_ms = kms_[0];
    *aux_ms = kms_[1];
    *main_n = kn_[0];
    *aux_n = kn_[1];
  }

 private:
  bool fail(const std::string& msg);
  bool hgx(int rc, const char* what);
  bool device();  // creates the device context on first use
  void release_device_objects();
  bool ensure_grid(double alphai_rad, double dnl_re, double dnl_im);
  bool run_device(double alpha_i_deg, double phi_deg, const double** d_img);
  bool make_grains(const HigStructure& s, GlibcRandom& rng, double phi_rad, std::vector<HigGrain>& out);
  // The grain sets of structure s: one per scaling sample and repetition class, in the reference's
  // loop order (a single set without distributions).
  bool grain_sets(const HigStructure& s, const std::vector<HigGrain>& grains, std::vector<GrainSet>& out);
  // d_I (+)= the DWBA intensity of one grain set.
  bool structure_intensity(const HigStructure& s, double alphai, const GrainSet& gs, bool accumulate, double* d_I);
  // The Fresnel coefficients for the particle layer at alphai (buffer 2; NULL for SAXS), cached.
  bool fresnel_coeffs(double alphai, double** d_fc);
  // The ensemble model of structure s for `grains` when its unit cell is analytic (false otherwise,
  // with the reason in err_ only when `report`).
  bool ensemble_model(const HigStructure& s, const std::vector<HigGrain>& grains, EnsembleBuild& b, bool report,
                      const vec3* scaling = nullptr, const vec3* repet = nullptr);
  // The normalised iratio weights of the structures (src/sim/hipgisaxs_main.cpp:1268-1282).
  std::vector<double> iratios() const;
  bool form_factor(const HigShape& shape, const Rot3& rot, const vec3& loc, double* d_out);
  bool shape_params(const HigShape& shape, hgx_shape_params& p, std::vector<double> store[8]);
  int mesh_of(const HigShape& shape);
  double* buffer(int slot, size_t bytes);

  int device_ = 0;
  std::unique_ptr<hgx_adapter::GpuContext> gpu_;
  HigInput in_;
  bool loaded_ = false, inited_ = false;
  int precision_ = HGX_PREC_FP64;
  int flags_ = 0;
  unsigned seed_ = 1;
  GlibcRandom rng_;
  double k0_ = 0.0;
  unsigned nrow_ = 0, ncol_ = 0;
  int64_t npix_ = 0;
  // the grid currently on the device: real (nbranch 1) or extended for (alphai, dnl)
  bool grid_ok_ = false, grid_ext_ = false;
  uint64_t grid_gen_ = 0;          // bumped at every grid rebuild
  std::vector<double> fc_key_;     // inputs of the Fresnel coefficients now in buffer 2
  double grid_ai_ = 0.0, grid_dn_[2] = {0.0, 0.0};
  // device buffers: 0 image, 1 structure image, 2 Fresnel coefficients, 3-5 per-grain form and
  // structure factors (tangent images in a gradient pass), 6 reference data, 7 mask, 8 a generation's
  // distances
  static constexpr int kBuffers = 9;
  void* buf_[kBuffers] = {};
  size_t buf_bytes_[kBuffers] = {};
  std::map<std::string, int> meshes_;
  bool fit_ready_ = false;
  int64_t fit_npix_ = 0;
  std::vector<std::vector<double>> history_;
  int64_t counts_[4] = {0, 0, 0, 0};
  int64_t paths_[3] = {0, 0, 0};
  double last_ms_ = 0.0;
  double kms_[2] = {0.0, 0.0};
  int64_t kn_[2] = {0, 0};
  std::string err_;
};

}  // namespace hig

// The handle behind the C ABI of include/hgx_hig.h.
struct hgx_hig {
  hig::HipGISAXS sim;
  std::string err;
  explicit hgx_hig(int device) : sim(device) {}
};

// Runs the body of a C entry point and turns any C++ exception into an error code, so that none
// crosses the extern "C" boundary.
template <class F>
int hgx_hig_guard(hgx_hig* h, F&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    if (h) h->err = "out of host memory";
    return HGX_ERR_OOM;
  } catch (const std::exception& e) {
    if (h) h->err = std::string("internal error: ") + e.what();
    return HGX_ERR_INVALID_ARG;
  } catch (...) {
    if (h) h->err = "internal error";
    return HGX_ERR_INVALID_ARG;
  }
}
