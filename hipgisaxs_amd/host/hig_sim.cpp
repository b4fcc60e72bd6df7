// hig_sim.cpp -- the .hig simulation driver on the hgx C ABI (see hig_sim.hpp).
#include "hig_sim.hpp"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>

namespace hig {

// ----- Rot3 (include/numerics/matrix.hpp:35-110) ---------------------------------------------------

Rot3::Rot3() : m{1, 0, 0, 0, 1, 0, 0, 0, 1} {}

Rot3::Rot3(int axis, double a) {
  const double c = std::cos(a), s = std::sin(a);
  const double x[9] = {1, 0, 0, 0, c, -s, 0, s, c};
  const double y[9] = {c, 0, -s, 0, 1, 0, s, 0, c};
  const double z[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
  const double* r = axis == 0 ? x : axis == 1 ? y : z;
  std::memcpy(m, r, sizeof(m));
}

Rot3 Rot3::operator*(const Rot3& rhs) const {
  Rot3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += m[3 * i + k] * rhs.m[3 * k + j];
      r.m[3 * i + j] = s;
    }
  return r;
}

std::array<double, 3> Rot3::operator*(const std::array<double, 3>& v) const {
  return {m[0] * v[0] + m[1] * v[1] + m[2] * v[2], m[3] * v[0] + m[4] * v[1] + m[5] * v[2],
          m[6] * v[0] + m[7] * v[1] + m[8] * v[2]};
}

namespace {

const double PI_ = 3.141592653589793;  // include/common/constants.hpp:29
const double SQRT_2PI_ = 2.506628;     // include/common/constants.hpp:30 (as truncated there)

// HipGISAXS::gaussian and ::cauchy (include/sim/hipgisaxs_main.hpp:151-172)
double gaussian_weight(double x, double mean, double sigma) {
  if (sigma < DBL_MIN) return (x - mean < DBL_MIN) ? 1.0 : 0.0;
  return (1.0 / (sigma * SQRT_2PI_) * std::exp(-(x - mean) * (x - mean) / (2 * sigma * sigma)));
}

double cauchy_weight(double x, double l, double s) {
  if (s > DBL_MIN) return 1.0 / (PI_ * s * (1.0 + ((x - l) / s) * ((x - l) / s)));
  return (x - l > DBL_MIN) ? 0.0 : 1.0;
}

bool same9(const double* a, const double* b) { return std::memcmp(a, b, 9 * sizeof(double)) == 0; }

// angle count of a scanned range, as run_all_gisaxs counts it (src/sim/hipgisaxs_main.cpp:298-312)
int scan_count(double mn, double mx, double step, bool alphai) {
  if (alphai) {
    if (mx < mn) mx = mn;
    if (mn == mx || step == 0) return 1;
  } else if (step == 0) {
    return 1;
  }
  int n = (int)((mx - mn) / step + 1);
  return n < 1 ? 1 : n;
}

}  // namespace

// Lattice::construct_vectors (src/model/structure.cpp:86-182): unit-scaled vectors of a lattice
// type and hkl; fco uses the class's ca_ = 1 and gamma_ = 0 (the file's caratio and abangle are
// never passed to it). The basis vector t is left out, as the CPU structure factor does.
bool lattice_vectors(const HigStructure& s, vec3& a, vec3& b, vec3& c, std::string& err) {
  if (s.abc_set) {
    a = s.a;
    b = s.b;
    c = s.c;
    return true;
  }
  const std::string& t = s.lattice_type;
  const std::string& hkl = s.hkl;
  if (t == "cubic" && hkl == "100") {
    a = {1, 0, 0}; b = {0, 1, 0}; c = {0, 0, 1};
  } else if (t == "cubic" && hkl == "110") {
    a = {1, 0, 0}; b = {0., 1.4142, 0}; c = {0., 0.7071, 0.7071};
  } else if (t == "bcc" && hkl == "100") {
    a = {1., 0., 0.}; b = {0., 1., 0.}; c = {0.5, 0.5, 0.5};
  } else if (t == "bcc" && hkl == "110") {
    a = {1.0, 0.0, 0}; b = {0.5, 0.7071, 0}; c = {0.0, 0.7071, 0.7071};
  } else if (t == "fcc" && hkl == "100") {
    a = {0.0, 0.5, 0.5}; b = {0.5, 0.0, 0.5}; c = {0.5, 0.5, 0.0};
  } else if (t == "fcc" && hkl == "111") {
    a = {1., 0., 0}; b = {0.5, 0.866, -0.}; c = {1., 0.5773, 0.866};
  } else if (t == "fco") {
    const double ca = 1.0, gamma = 0.0;
    a = {1, 0, 0}; b = {0.5, std::tan(gamma) / 2.0, 0}; c = {0, 0, ca};
  } else if (t == "hcp") {
    a = {1, 0, 0}; b = {0.5, 0.866, 0}; c = {0.5, 0.2887, 0.8165};
  } else if (t == "hex") {
    a = {1, 0, 0}; b = {0.5, std::sqrt(3.0) / 2.0, 0}; c = {0, 0, 1};
  } else {
    err = "invalid lattice type and hkl combination '" + t + "', '" + hkl + "'";
    return false;
  }
  return true;
}

// ----- GlibcRandom ---------------------------------------------------------------------------------

void GlibcRandom::reseed(unsigned seed) {
  // r[0] is the seed as a signed 32-bit word, and the LCG runs on signed 32-bit words (so a seed of 2^31
  // or more starts from its negative value, as the C library's does)
  int32_t init[34];
  init[0] = (int32_t)(seed == 0 ? 1 : seed);
  for (int i = 1; i < 31; ++i) {
    // 16807 * r mod (2^31 - 1) by Schrage's decomposition (no 32-bit overflow)
    const int32_t hi = init[i - 1] / 127773, lo = init[i - 1] % 127773;
    int32_t w = 16807 * lo - 2836 * hi;
    if (w < 0) w += 2147483647;
    init[i] = w;
  }
  for (int i = 31; i < 34; ++i) init[i] = init[i - 31];
  for (int i = 0; i < 34; ++i) r_[i] = (uint32_t)init[i];
  i_ = 0;
  for (int k = 0; k < 310; ++k) next();
}

int GlibcRandom::next() {
  // ring slot i_ holds r[n-34]; r[n] = r[n-31] + r[n-3] replaces it
  uint32_t v = r_[(i_ + 3) % 34] + r_[(i_ + 31) % 34];
  r_[i_] = v;
  i_ = (i_ + 1) % 34;
  return (int)(v >> 1);
}

// ----- HipGISAXS -----------------------------------------------------------------------------------

HipGISAXS::HipGISAXS(int device) : device_(device) {}

HipGISAXS::~HipGISAXS() { release_device_objects(); }

bool HipGISAXS::device() {
  if (gpu_ && gpu_->ok()) return true;
  gpu_.reset(new hgx_adapter::GpuContext(device_));
  if (!gpu_->ok()) {
    gpu_.reset();
    return fail("could not create the device context on device " + std::to_string(device_));
  }
  return true;
}

void HipGISAXS::release_device_objects() {
  fc_key_.clear();
  if (gpu_ && gpu_->ok()) {
    for (auto& m : meshes_) hgx_mesh_release(gpu_->get(), m.second);
    for (int i = 0; i < kBuffers; ++i)
      if (buf_[i]) hgx_device_free(gpu_->get(), buf_[i]);
  }
  meshes_.clear();
  for (int i = 0; i < kBuffers; ++i) {
    buf_[i] = nullptr;
    buf_bytes_[i] = 0;
  }
}

bool HipGISAXS::fail(const std::string& msg) {
  err_ = msg;
  return false;
}

bool HipGISAXS::hgx(int rc, const char* what) {
  if (rc == HGX_OK) return true;
  err_ = std::string(what) + " failed (" + std::to_string(rc) + "): " + (gpu_ ? hgx_last_error(gpu_->get()) : "");
  return false;
}

double* HipGISAXS::buffer(int slot, size_t bytes) {
  if (bytes <= buf_bytes_[slot]) return static_cast<double*>(buf_[slot]);
  if (buf_[slot]) hgx_device_free(gpu_->get(), buf_[slot]);
  buf_[slot] = nullptr;
  buf_bytes_[slot] = 0;
  if (hgx_device_alloc(gpu_->get(), bytes, &buf_[slot]) != HGX_OK) return nullptr;
  buf_bytes_[slot] = bytes;
  return static_cast<double*>(buf_[slot]);
}

void HipGISAXS::options(int precision, int flags, unsigned seed) {
  precision_ = precision;
  flags_ = flags;
  seed_ = seed;
  rng_.reseed(seed_);
}

bool HipGISAXS::construct_input(const std::string& path) {
  loaded_ = inited_ = fit_ready_ = false;
  history_.clear();
  if (gpu_ && gpu_->ok())
    for (auto& m : meshes_) hgx_mesh_release(gpu_->get(), m.second);
  meshes_.clear();
  if (!in_.construct_input(path)) return fail(in_.error());
  loaded_ = true;
  rng_.reseed(seed_);
  return true;
}

bool HipGISAXS::construct_input_text(const std::string& text, const std::string& base_dir) {
  loaded_ = inited_ = fit_ready_ = false;
  history_.clear();
  if (gpu_ && gpu_->ok())
    for (auto& m : meshes_) hgx_mesh_release(gpu_->get(), m.second);
  meshes_.clear();
  if (!in_.construct_input_text(text, base_dir)) return fail(in_.error());
  loaded_ = true;
  rng_.reseed(seed_);
  return true;
}

bool HipGISAXS::prepare() {
  if (!loaded_) return fail("no input has been read");
  // photon energy -> wavelength -> k0 (src/sim/hipgisaxs_main.cpp:139-158)
  double photon = in_.scattering.photon, lambda;
  if (in_.scattering.unit == "ev") {
    photon = photon / 1000;
    lambda = 1.23984 / photon;
  } else if (in_.scattering.unit == "kev") {
    lambda = 1.23984 / photon;
  } else {
    return fail("photon energy is not given in 'ev' or 'kev'");
  }
  if (!(lambda > 0) || !std::isfinite(lambda)) return fail("photon energy must be positive");
  k0_ = 2 * PI_ / lambda;
  const HigCompute& cp = in_.compute;
  if (cp.region != "qspace")
    return fail("output region '" + cp.region + "' is not supported (the reference implements qspace only)");
  ncol_ = (unsigned)cp.resolution[0];
  nrow_ = (unsigned)cp.resolution[1];
  if (ncol_ < 2 || nrow_ < 2) return fail("resolution must be at least [2 2]");
  npix_ = (int64_t)ncol_ * nrow_;
  return true;
}

bool HipGISAXS::init() {
  if (inited_) return true;
  if (!prepare() || !device()) return false;
  grid_ok_ = false;  // a new input: rebuild the grid
  if (!ensure_grid(in_.scattering.alphai_min * PI_ / 180, 0.0, 0.0)) return false;
  inited_ = true;
  return true;
}

// The q-grid of the output region is built at the smallest incidence angle (HipGISAXS::init,
// src/sim/hipgisaxs_main.cpp:171-176); a GISAXS run extends its q_z for the run's incidence angle and
// the particle layer (layer_qgrid_qz, :1548-1556). alphai < 0 requests the plain (SAXS) grid.
bool HipGISAXS::ensure_grid(double alphai, double dnl_re, double dnl_im) {
  const bool want_ext = alphai >= 0 && in_.scattering.expt == "gisaxs";
  if (grid_ok_ && !want_ext && !grid_ext_) return true;
  if (grid_ok_ && want_ext && grid_ext_ && grid_ai_ == alphai && grid_dn_[0] == dnl_re && grid_dn_[1] == dnl_im)
    return true;
  grid_ok_ = false;
  const HigCompute& cp = in_.compute;
  double ai_min = in_.scattering.alphai_min * PI_ / 180;
  if (!hgx(hgx_qgrid_qspace(gpu_->get(), (int)ncol_, (int)nrow_, cp.minpoint[0], cp.minpoint[1], cp.maxpoint[0],
                            cp.maxpoint[1], ai_min, k0_),
           "hgx_qgrid_qspace"))
    return false;
  grid_ext_ = false;
  ++grid_gen_;  // coefficients computed on an earlier grid are stale
  if (want_ext && !hgx(hgx_qz_extended(gpu_->get(), k0_, alphai, dnl_re, dnl_im), "hgx_qz_extended")) return false;
  grid_ok_ = true;
  if (!want_ext) return true;
  grid_ext_ = true;
  grid_ai_ = alphai;
  grid_dn_[0] = dnl_re;
  grid_dn_[1] = dnl_im;
  return true;
}

std::vector<std::array<double, 3>> HipGISAXS::configurations() const {
  const HigScattering& sc = in_.scattering;
  double amax = sc.alphai_max < sc.alphai_min ? sc.alphai_min : sc.alphai_max;
  int na = scan_count(sc.alphai_min, amax, sc.alphai_step, true);
  int np = scan_count(sc.phi_min, sc.phi_max, sc.phi_step, false);
  int nt = scan_count(sc.tilt_min, sc.tilt_max, sc.tilt_step, false);
  std::vector<std::array<double, 3>> out;
  // angles advance by accumulation, as the reference's loops do (:350-351, :383, :410)
  double a = sc.alphai_min;
  for (int i = 0; i < na; ++i, a += sc.alphai_step) {
    double p = sc.phi_min;
    for (int j = 0; j < np; ++j, p += sc.phi_step) {
      double t = sc.tilt_min;
      for (int k = 0; k < nt; ++k, t += sc.tilt_step) out.push_back({a, p, t});
    }
  }
  return out;
}

bool HipGISAXS::compute_gisaxs(std::vector<double>& img) {
  if (!loaded_) return fail("no input");
  if (configurations().size() != 1)
    return fail("currently you can simulate only for single alpha_i, phi and tilt angles");
  const HigScattering& sc = in_.scattering;
  return run_gisaxs(sc.alphai_min, sc.phi_min, sc.tilt_min, img);
}

// spatial_distribution and orientation_distribution (src/sim/hipgisaxs_main.cpp:1702-2086) and the
// grain rotation and centre of run_gisaxs (:925-958).
bool HipGISAXS::make_grains(const HigStructure& s, GlibcRandom& rng, double phi, std::vector<HigGrain>& out) {
  out.clear();
  if (s.distribution != "regular")
    return fail("structure '" + s.key + "': ensemble distribution '" + s.distribution +
                "' is not supported (regular only; the reference's random placement reseeds rand() from the clock)");
  for (double m : s.maxgrains)
    if (!(m >= 1 && m <= 1e6 && m == std::floor(m)))
      return fail("structure '" + s.key + "': maxgrains must be positive integers up to 1e6");
  const double ngd = s.maxgrains[0] * s.maxgrains[1] * s.maxgrains[2];
  if (ngd > 1000000) return fail("structure '" + s.key + "': grain count out of range (1 to 1e6)");
  const int nd[3] = {(int)s.maxgrains[0], (int)s.maxgrains[1], (int)s.maxgrains[2]};
  const int ng = nd[0] * nd[1] * nd[2];
  // grain positions: cell_ of the reference is never set (zero), so the spacing alone separates grains
  const double tz = 0.0;
  std::vector<double> dx(nd[0]), dy(nd[1]), dz(nd[2]);
  {
    double val = 0, step = s.spacing[0], mx = s.spacing[0] * (nd[0] - 1);
    for (int i = 0; i < nd[0]; ++i, val += step) dx[i] = val - mx / 2;
    val = 0, step = s.spacing[1], mx = s.spacing[1] * (nd[1] - 1);
    for (int i = 0; i < nd[1]; ++i, val += step) dy[i] = val - mx / 2;
    val = 0, step = s.spacing[2];
    for (int i = 0; i < nd[2]; ++i, val += step) dz[i] = val + tz;
  }
  // orientations: nn[k][g] (radians) and weights
  std::vector<double> nn(3 * (size_t)ng, 0.0), wght(3 * (size_t)ng, 1.0);
  const std::string& st = s.orient_stat;
  HigRotation rot[3];
  int axis[3];
  for (int k = 0; k < 3; ++k) {
    rot[k] = s.rot[k];
    if (rot[k].axis == 'n') {  // Structure::rotation_rotN returns (0, 0, 0) for axis 'n'
      rot[k].axis = 'x';
      rot[k].amin = rot[k].amax = 0.0;
    }
    axis[k] = rot[k].axis - 'x';
  }
  bool gauss = false;
  if (st == "single") {
    for (int k = 0; k < 3; ++k)
      for (int g = 0; g < ng; ++g) nn[k * ng + g] = rot[k].amin * PI_ / 180;
  } else if (st == "random") {
    for (size_t x = 0; x < nn.size(); ++x) nn[x] = (double(rng.next()) / GlibcRandom::kRandMax) * 2 * PI_;
  } else if (st == "range") {
    for (int k = 0; k < 3; ++k) {
      double d = std::fabs(rot[k].amax - rot[k].amin);
      for (int g = 0; g < ng; ++g)
        nn[k * ng + g] = (rot[k].amin + (double(rng.next()) / GlibcRandom::kRandMax) * d) * PI_ / 180;
    }
  } else if (st == "gaussian" || st == "normal" || st == "cauchy") {
    // run_gisaxs multiplies the grain weights in for "gaussian" and "cauchy" only; "normal" samples the
    // same angles and weights but leaves the grains unweighted (src/sim/hipgisaxs_main.cpp:934-938)
    gauss = st != "normal";
    const bool cauchy = st == "cauchy";
    for (int k = 0; k < 3; ++k) {
      double d = std::fabs(rot[k].amax - rot[k].amin);
      d = cauchy ? (ng > 1 ? d / (ng - 1) : 0.0) : d;  // gaussian divides by ng below, cauchy by ng - 1
      if (d > DBL_MIN) {
        if (!cauchy) d /= ng;
        for (int g = 0; g < ng; ++g) {
          double t = rot[k].amin + g * d;
          nn[k * ng + g] = t * PI_ / 180;
          wght[k * ng + g] = cauchy ? cauchy_weight(t, rot[k].location, rot[k].scale)
                                    : gaussian_weight(t, rot[k].mean, rot[k].sd);
        }
      } else {
        for (int g = 0; g < ng; ++g) {
          nn[k * ng + g] = rot[k].amin * PI_ / 180;
          wght[k * ng + g] = 1;
        }
      }
    }
  } else {
    return fail("structure '" + s.key + "': orientation statistic '" + st + "' is not supported");
  }
  vec3 tv = s.transvec;  // minus the reference's never-initialised single_layer_thickness_, taken as 0
  Rot3 rphi(2, phi);
  out.resize(ng);
  const int nyz = nd[1] * nd[2];
  for (int g = 0; g < ng; ++g) {
    Rot3 r1(axis[0], nn[g]), r2(axis[1], nn[ng + g]), r3(axis[2], nn[2 * ng + g]);
    HigGrain& G = out[g];
    G.rot = rphi * r3 * r2 * r1;
    // d1[nyz x + yz] = dx[x], d2[nyz x + nz y + z] = dy[y], d3[nz xy + z] = dz[z]
    int x = g / nyz, y = (g % nyz) / nd[2], z = g % nd[2];
    std::array<double, 3> dd{dx[x], dy[y], dz[z]};
    std::array<double, 3> c = G.rot * dd;
    G.center[0] = c[0] + tv[0];
    G.center[1] = c[1] + tv[1];
    G.center[2] = c[2] + tv[2];
    G.weight = gauss ? wght[g] * wght[ng + g] * wght[2 * ng + g] : 1.0;
  }
  return true;
}

bool HipGISAXS::grains(const std::string& key, std::vector<HigGrain>& out) {
  if (!loaded_) return fail("no input");
  if (!in_.structures.count(key)) return fail("no structure '" + key + "'");
  GlibcRandom rng = rng_;  // the sampler's state is left as it is
  double phi = in_.scattering.phi_min * PI_ / 180;
  for (const auto& s : in_.structures) {
    if (!make_grains(s.second, rng, phi, out)) return false;
    if (s.first == key) return true;
  }
  return false;
}

// Distribution nodes of an analytic shape's parameters in the order hgx_ff_analytic takes them
// (the parameter switches of src/ff/ff_ana_<shape>.cpp; the cube reads its edge from "height",
// ff_ana_cube.cpp:54-56; a box "edge" sets all three sides unless a side is given).
bool HipGISAXS::shape_params(const HigShape& sh, hgx_shape_params& p, std::vector<double> store[8]) {
  std::memset(&p, 0, sizeof(p));
  std::vector<std::vector<std::string>> dims;
  switch (sh.shape) {
    case HGX_SHAPE_SPHERE:
      // the reference GPU path cuts spheres that carry a height (ff_ana_sphere.cpp:53-57, GPU :180-188)
      if ((flags_ & HGX_HIG_TRUNCATED_SPHERE) && sh.params.count("height")) {
        dims = {{"radius"}, {"height"}};
        p.flags |= HGX_SHAPE_TRUNCATED_SPHERE;
      } else {
        dims = {{"radius"}};
      }
      break;
    case HGX_SHAPE_CYLINDER: dims = {{"radius"}, {"height"}}; break;
    case HGX_SHAPE_BOX: dims = {{"xsize", "edge"}, {"ysize", "edge"}, {"height", "edge"}}; break;
    case HGX_SHAPE_CUBE: dims = {{"height"}}; break;
    case HGX_SHAPE_PRISM3:
    case HGX_SHAPE_PRISM6: dims = {{"edge"}, {"height"}}; break;
    case HGX_SHAPE_PYRAMID: dims = {{"xsize"}, {"ysize"}, {"height"}, {"baseangle"}}; break;
    default: return fail("shape '" + sh.key + "' (" + sh.name + ") has no analytic form factor");
  }
  for (size_t d = 0; d < dims.size(); ++d) {
    const HigShapeParam* prm = nullptr;
    for (const auto& name : dims[d]) {
      auto it = sh.params.find(name);
      if (it != sh.params.end()) {
        prm = &it->second;
        break;
      }
    }
    if (!prm) return fail("shape '" + sh.key + "' (" + sh.name + ") needs a '" + dims[d][0] + "' parameter");
    if (prm->stat == 2 && !(prm->p2 > 0))
      return fail("shape '" + sh.key + "': gaussian parameter '" + prm->type + "' needs p2 (deviation) > 0");
    int cap = 4 * prm->nvalues + 8;
    store[2 * d].resize(cap);
    store[2 * d + 1].resize(cap);
    // the reference's node grid and weights, so that an input gives the reference's image
    int n = hgx_param_distribution(prm->min, prm->max, prm->nvalues, prm->stat, prm->p1, prm->p2,
                                   HGX_DIST_REFERENCE, store[2 * d].data(), store[2 * d + 1].data(), cap);
    if (n < 1) return fail("shape '" + sh.key + "': bad distribution for parameter '" + prm->type + "'");
    if (n > HGX_MAX_NODES)
      return fail("shape '" + sh.key + "': parameter '" + prm->type + "' has more than " +
                  std::to_string(HGX_MAX_NODES) + " nodes");
    p.n[d] = n;
    p.val[d] = store[2 * d].data();
    p.wt[d] = store[2 * d + 1].data();
  }
  return true;
}

int HipGISAXS::mesh_of(const HigShape& sh) {
  auto it = meshes_.find(sh.key);
  if (it != meshes_.end()) return it->second;
  std::vector<double> rec;
  std::string err;
  int kind;
  if (sh.file.size() > 4 && sh.file.compare(sh.file.size() - 4, 4, ".obj") == 0) {
    if (!read_obj_mesh(sh.file, rec, err)) return fail(err), -1;
    kind = HGX_MESH_EXACT9;
  } else if (sh.file.size() > 4 && sh.file.compare(sh.file.size() - 4, 4, ".dat") == 0) {
    if (!read_raw_mesh(sh.file, rec, err)) return fail(err), -1;
    kind = HGX_MESH_APPROX8;
  } else {
    return fail("mesh file '" + sh.file + "': only .obj and .dat meshes are supported"), -1;
  }
  int id = -1;
  int64_t n = (int64_t)(rec.size() / (kind == HGX_MESH_EXACT9 ? 9 : 8));
  if (!hgx(hgx_mesh_upload(gpu_->get(), kind, n, rec.data(), &id), "hgx_mesh_upload")) return -1;
  meshes_[sh.key] = id;
  return id;
}

// FormFactor::compute_form_factor (src/ff/ff.cpp:32-61): analytic shapes with the location phase,
// mesh shapes by the approximate (centroid) method of NumericFormFactor::compute
// (src/ff/ff_num.cpp:187-284) under the run's precision.
bool HipGISAXS::form_factor(const HigShape& sh, const Rot3& rot, const vec3& loc, double* d_out) {
  ++counts_[2];
  if (sh.shape == -1) {
    int id = mesh_of(sh);
    if (id < 0) return false;
    const double* tv = (flags_ & HGX_HIG_MESH_NO_TRANSLATION) ? nullptr : loc.data();
    int path = HGX_PATH_PAIR;
    if (!hgx(hgx_ff_numeric_path(gpu_->get(), id, HGX_FF_APPROX, precision_, 0, &path, nullptr), "hgx_ff_numeric_path"))
      return false;
    ++paths_[path == HGX_PATH_PRODUCT ? 1 : path == HGX_PATH_EWALD ? 2 : 0];
    return hgx(hgx_ff_numeric(gpu_->get(), id, HGX_FF_APPROX, precision_, 0, rot.data(), tv, d_out), "hgx_ff_numeric");
  }
  hgx_shape_params p;
  std::vector<double> store[8];
  if (!shape_params(sh, p, store)) return false;
  return hgx(hgx_ff_analytic(gpu_->get(), sh.shape, &p, precision_, rot.data(), loc.data(), d_out),
             "hgx_ff_analytic");
}

// One structure's grain sum into d_I (written, not accumulated): the body of run_gisaxs's structure
// loop for the nGnE correlation (src/sim/hipgisaxs_main.cpp:775-1222).
void EnsembleBuild::relink() {
  m.nelem = (int)el.size();
  m.elem = el.data();
  for (size_t e = 0; e < el.size(); ++e) {
    for (int d = 0; d < 4; ++d) {
      el[e].params.val[d] = el[e].params.n[d] > 0 ? stores[8 * e + 2 * d].data() : nullptr;
      el[e].params.wt[d] = el[e].params.n[d] > 0 ? stores[8 * e + 2 * d + 1].data() : nullptr;
    }
    el[e].loc = locs[e].data();
  }
  m.ngrains = (int)wts.size();
  m.rot = rots.data();
  m.center = cen.data();
  m.weight = wts.data();
}

std::vector<double> HipGISAXS::iratios() const {
  std::vector<double> r;
  double sum = 0.0;
  for (const auto& s : in_.structures) {
    r.push_back(s.second.iratio);
    sum += s.second.iratio;
  }
  if (sum != 1.0)
    for (double& x : r) x /= sum;
  return r;
}

bool HipGISAXS::ensemble_model(const HigStructure& s, const std::vector<HigGrain>& grains, EnsembleBuild& b,
                               bool report, const vec3* scaling, const vec3* repet) {
  const vec3& sc = scaling ? *scaling : s.scaling;
  const vec3& rp = repet ? *repet : s.repetition;
  auto no = [&](const std::string& why) { return report ? fail(why) : false; };
  auto uit = in_.unitcells.find(s.unitcell_key);
  if (uit == in_.unitcells.end()) return no("structure '" + s.key + "': no unitcell '" + s.unitcell_key + "'");
  const HigUnitcell& uc = uit->second;
  for (const auto& e : uc.elements) {
    auto sh = in_.shapes.find(e.first);
    if (sh == in_.shapes.end()) return no("unitcell '" + uc.key + "': no shape '" + e.first + "'");
    if (sh->second.shape < 0) return no("shape '" + e.first + "' is not analytic");
  }
  vec3 a, bb, c;
  std::string err;
  if (!lattice_vectors(s, a, bb, c, err)) return no("structure '" + s.key + "': " + err);
  const int ne = (int)uc.elements.size(), ng = (int)grains.size();
  if (ne < 1 || ne > 64 || ng < 1) return no("structure '" + s.key + "': unit cell or grains out of range");
  // the particle layer is the vacuum (structure_intensity)
  const double dnl[2] = {0.0, 0.0};
  b.el.assign(ne, hgx_ensemble_element{});
  b.stores.assign(8 * (size_t)ne, {});
  b.locs.assign(ne, {});
  int e = 0;
  for (const auto& it : uc.elements) {
    const HigShape& sh = in_.shapes.at(it.first);
    hgx_ensemble_element& E = b.el[e];
    std::memset(&E, 0, sizeof(E));
    if (!shape_params(sh, E.params, &b.stores[8 * (size_t)e])) return report ? false : (err_.clear(), false);
    E.shape = sh.shape;
    Rot3 re = Rot3(0, sh.xrot) * Rot3(1, sh.yrot) * Rot3(2, sh.zrot);
    std::memcpy(E.rot, re.data(), sizeof(E.rot));
    for (const auto& l : it.second) b.locs[e].insert(b.locs[e].end(), l.begin(), l.end());
    E.nloc = (int)it.second.size();
    E.dn2_re = dnl[0] - 2 * sh.delta;
    E.dn2_im = dnl[1] - 2 * sh.beta;
    ++e;
  }
  std::memset(&b.m, 0, sizeof(b.m));
  // StructureFactor::compute_structure_factor scales the vectors (src/sf/sf.cpp:165-171)
  for (int i = 0; i < 3; ++i) {
    b.m.la[i] = a[i] * sc[0];
    b.m.lb[i] = bb[i] * sc[1];
    b.m.lc[i] = c[i] * sc[2];
    b.m.repet[i] = rp[i];
  }
  b.m.dw_sigma = 0.0;
  b.m.sf_flags = (flags_ & HGX_HIG_SF_COMPAT_LB2) ? HGX_SF_COMPAT_LB2 : 0;
  b.rots.assign(9 * (size_t)ng, 0.0);
  b.cen.assign(3 * (size_t)ng, 0.0);
  b.wts.assign(ng, 0.0);
  for (int g = 0; g < ng; ++g) {
    std::memcpy(&b.rots[9 * (size_t)g], grains[g].rot.data(), 9 * sizeof(double));
    std::memcpy(&b.cen[3 * (size_t)g], grains[g].center, 3 * sizeof(double));
    b.wts[g] = grains[g].weight;
  }
  b.relink();
  return true;
}

// The Fresnel coefficients of the layer stack for the particle layer at alphai, in buffer 2 (SAXS:
// none), recomputed only when the stack, k0, the angle or the grid changed.
bool HipGISAXS::fresnel_coeffs(double alphai, double** d_out) {
  *d_out = nullptr;
  if (in_.scattering.expt == "gisaxs") {
    // MultiLayer::init (src/model/multilayer.cpp:29-59): layers of order >= 0 top to bottom, then the
    // substrate (order -1, or an empty layer when the input has none)
    std::vector<double> de, be, th;
    for (const auto& l : in_.layers) {
      if (l.first < 0) continue;
      if (l.first == 0 && l.second.key == "vacuum" && l.second.delta == 0 && l.second.beta == 0 &&
          l.second.thickness == 0)
        continue;  // the implicit vacuum: a zero-thickness vacuum layer leaves the coefficients unchanged
      de.push_back(l.second.delta);
      be.push_back(l.second.beta);
      th.push_back(l.second.thickness);
    }
    auto sub = in_.layers.find(-1);
    de.push_back(sub != in_.layers.end() ? sub->second.delta : 0.0);
    be.push_back(sub != in_.layers.end() ? sub->second.beta : 0.0);
    th.push_back(sub != in_.layers.end() ? sub->second.thickness : 0.0);
    const int64_t nq = npix_ * 4;
    const int order = 0;  // the particle layer (see structure_intensity)
    double* d_fc = buffer(2, (size_t)nq * 16);
    if (!d_fc) return fail("device allocation failed");
    // the coefficients depend on the stack, k0, the incidence angle and the grid's rows: reuse them while
    // none of these changed (structures of one run, successive fit evaluations)
    std::vector<double> key = de;
    key.insert(key.end(), be.begin(), be.end());
    key.insert(key.end(), th.begin(), th.end());
    key.insert(key.end(), {k0_, alphai, (double)nq, (double)grid_gen_, (double)(uintptr_t)d_fc});
    if (key != fc_key_) {
      fc_key_.clear();
      if (!hgx(hgx_fresnel(gpu_->get(), (int)de.size(), de.data(), be.data(), th.data(), order, k0_, alphai, d_fc),
               "hgx_fresnel"))
        return false;
      fc_key_ = key;
    }
    *d_out = d_fc;
  }
  return true;
}

namespace {

// HipGISAXS::gaussian / cauchy (include/sim/hipgisaxs_main.hpp:151-172), with their constants
double scaling_pdf(int stat, double x, double mean, double sd) {
  const double kTiny = DBL_MIN, kSqrt2Pi = 2.506628, kPi = 3.141592653589793;
  if (stat == HigStructure::kStatGaussian) {
    if (sd < kTiny) return (x - mean < kTiny) ? 1.0 : 0.0;
    return 1.0 / (sd * kSqrt2Pi) * std::exp(-(x - mean) * (x - mean) / (2 * sd * sd));
  }
  if (sd > kTiny) return 1.0 / (kPi * sd * (1.0 + ((x - mean) / sd) * ((x - mean) / sd)));
  return (x - mean > kTiny) ? 0.0 : 1.0;
}

}  // namespace

bool HipGISAXS::grain_sets(const HigStructure& s, const std::vector<HigGrain>& grains, std::vector<GrainSet>& out) {
  out.clear();
  const int ng = (int)grains.size();
  // construct_scaling_distribution (src/sim/hipgisaxs_main.cpp:819-841, 2179-2226): each axis gives
  // its mean (no statistic) or max_nvals samples in [mean - 4 sd, mean + 4 sd], one per cell of
  // width 8 sd / (max_nvals + 1) at a uniform offset from one generator seeded 0, weighted by the pdf;
  // an axis whose width is zero counts as one sample in max_nvals
  int nvals[3];
  for (int i = 0; i < 3; ++i)
    nvals[i] = s.scale_stat[i] != HigStructure::kStatNone ? (s.scale_sd[i] == 0 ? 1 : s.scale_nvals[i]) : 1;
  const int maxn = std::max(nvals[0], std::max(nvals[1], nvals[2]));
  MtUniform rng(0);
  std::vector<double> smp, wts;
  for (int i = 0; i < 3; ++i) {
    const int st = s.scale_stat[i];
    if (st == HigStructure::kStatNone) {
      smp.push_back(s.scaling[i]);
      wts.push_back(1.0);
    } else if (st == HigStructure::kStatGaussian || st == HigStructure::kStatCauchy) {
      const double lo = s.scaling[i] - 4 * s.scale_sd[i], dx = 8 * s.scale_sd[i] / (maxn + 1);
      for (int j = 0; j < maxn; ++j) {
        const double t = lo + j * dx + rng.next() * dx;
        smp.push_back(t);
        wts.push_back(scaling_pdf(st, t, s.scaling[i], s.scale_sd[i]));
      }
    } else {
      return fail("structure '" + s.key + "': unsupported distribution for scaling (none, gaussian or cauchy)");
    }
  }
  // construct_repetition_distribution (:843-857, 2089-2176): per axis the minimum, or n values
  // min + u (max - min) truncated, u from a generator seeded 0 for each axis; the axes are zipped to the
  // shortest. n is the size of the scaling sample list (3 values even without a distribution), or the
  // grain count for a list of one.
  std::vector<vec3> reps;
  if (s.repetition_dist) {
    const int nr = smp.size() > 1 ? (int)smp.size() : ng;
    std::vector<double> vals[3];
    for (int ax = 0; ax < 3; ++ax) {
      if (s.rep_stat[ax] == 0) {
        vals[ax].push_back(s.rep_min[ax]);
      } else if (s.rep_stat[ax] == 1) {
        if (s.rep_max[ax] < s.rep_min[ax])
          return fail("structure '" + s.key + "': repetition range with max < min");
        MtUniform r(0);
        for (int k = 0; k < nr; ++k)
          vals[ax].push_back(std::floor(s.rep_min[ax] + r.next() * (s.rep_max[ax] - s.rep_min[ax])));
      } else {
        return fail("structure '" + s.key + "': grain repetition distributions take stat none or range per axis");
      }
    }
    const size_t n = std::min(vals[0].size(), std::min(vals[1].size(), vals[2].size()));
    for (size_t k = 0; k < n; ++k) reps.push_back(vec3{vals[0][k], vals[1][k], vals[2][k]});
  } else {
    reps.push_back(s.repetition);
  }
  // the loop (:859-868, 1031-1061): grain scaling from the sample list at i, i + n, i + 2n (n = list
  // size / 3) when some width is positive, else its first three entries, with weight entry i (or 0);
  // repetitions by grain index modulo their count
  const int nsc = (int)smp.size() / 3, nrep = (int)reps.size();
  int nrs = 0;
  if (nsc == nrep) nrs = nsc;
  else if (nsc == 1) nrs = nrep;
  else if (nrep == 1) nrs = nsc;
  else return fail("structure '" + s.key + "': scaling and repetition distributions are not aligned");
  bool isdist = false;
  for (int i = 0; i < 3; ++i) isdist = isdist || s.scale_sd[i] > 0;
  const int nclass = s.repetition_dist ? nrep : 1;
  for (int is = 0; is < nrs; ++is) {
    vec3 sc;
    double w;
    if (isdist) {
      if (is + 2 * nsc >= (int)smp.size())
        return fail("structure '" + s.key + "': the reference reads past its scaling samples here");
      sc = {smp[is], smp[is + nsc], smp[is + 2 * nsc]};
      w = wts[is];
    } else {
      sc = {smp[0], smp[1], smp[2]};
      w = wts[0];
    }
    for (int r = 0; r < nclass; ++r) {
      GrainSet gs;
      gs.scaling = sc;
      gs.repet = reps[r];
      for (int g = 0; g < ng; ++g) {
        if (s.repetition_dist && g % nrep != r) continue;
        gs.grains.push_back(grains[g]);
        gs.grains.back().weight *= w;
      }
      if (!gs.grains.empty()) out.push_back(std::move(gs));
    }
  }
  return true;
}

bool HipGISAXS::structure_intensity(const HigStructure& s, double alphai, const GrainSet& gs, bool accumulate,
                                    double* d_I) {
  const std::vector<HigGrain>& grains = gs.grains;
  auto uit = in_.unitcells.find(s.unitcell_key);
  if (uit == in_.unitcells.end()) return fail("structure '" + s.key + "': no unitcell '" + s.unitcell_key + "'");
  const HigUnitcell& uc = uit->second;
  for (const auto& e : uc.elements)
    if (!in_.shapes.count(e.first)) return fail("unitcell '" + uc.key + "': no shape '" + e.first + "'");
  const bool gisaxs = in_.scattering.expt == "gisaxs";
  // The particle layer: the reference reads Structure::layer_order(), which no code path sets, so it
  // is always 0 -- the vacuum (structure.cpp:358, hipgisaxs_main.cpp:780).
  const double dnl[2] = {0.0, 0.0};
  if (!ensure_grid(gisaxs ? alphai : -1.0, dnl[0], dnl[1])) return false;
  const int nbranch = gisaxs ? 4 : 1;
  const int64_t nq = npix_ * nbranch;
  double* d_fc = nullptr;
  if (!fresnel_coeffs(alphai, &d_fc)) return false;
  vec3 a, b, c;
  std::string err;
  if (!lattice_vectors(s, a, b, c, err)) return fail("structure '" + s.key + "': " + err);
  // StructureFactor::compute_structure_factor scales the vectors (src/sf/sf.cpp:165-171)
  double la[3], lb[3], lc[3];
  for (int i = 0; i < 3; ++i) {
    la[i] = a[i] * gs.scaling[0];
    lb[i] = b[i] * gs.scaling[1];
    lc[i] = c[i] * gs.scaling[2];
  }
  const double repet[3] = {gs.repet[0], gs.repet[1], gs.repet[2]};
  const int sf_flags = (flags_ & HGX_HIG_SF_COMPAT_LB2) ? HGX_SF_COMPAT_LB2 : 0;
  const int ng = (int)grains.size();

  // the fused grain-ensemble kernel: one analytic sphere or cylinder at the cell origin, unrotated
  // against the grain, every grain centred alike
  bool fuse = !(flags_ & HGX_HIG_NO_FUSE) && uc.elements.size() == 1 && uc.elements.begin()->second.size() == 1;
  const HigShape* sh0 = fuse ? &in_.shapes.at(uc.elements.begin()->first) : nullptr;
  if (fuse) {
    const vec3& l0 = uc.elements.begin()->second[0];
    fuse = (sh0->shape == HGX_SHAPE_SPHERE || sh0->shape == HGX_SHAPE_CYLINDER) && l0[0] == 0 && l0[1] == 0 &&
           l0[2] == 0 && sh0->xrot == 0 && sh0->yrot == 0 && sh0->zrot == 0;
    for (int g = 1; fuse && g < ng; ++g)
      fuse = std::memcmp(grains[g].center, grains[0].center, sizeof(grains[0].center)) == 0;
  }
  if (fuse) {
    hgx_lattice_model m;
    std::memset(&m, 0, sizeof(m));
    std::vector<double> store[8];
    if (!shape_params(*sh0, m.params, store)) return false;
    m.shape = sh0->shape;
    for (int i = 0; i < 3; ++i) {
      m.la[i] = la[i];
      m.lb[i] = lb[i];
      m.lc[i] = lc[i];
      m.repet[i] = repet[i];
      m.center[i] = grains[0].center[i];
    }
    m.dw_sigma = 0.0;
    // contrast (1 - n^2)_layer - (1 - n^2)_shape, one_minus_n2 = 2 delta + 2 i beta
    m.dn2_re = dnl[0] - 2 * sh0->delta;
    m.dn2_im = dnl[1] - 2 * sh0->beta;
    m.sf_flags = sf_flags;
    m.ngrains = ng;
    std::vector<double> rots(9 * (size_t)ng), wts(ng);
    for (int g = 0; g < ng; ++g) {
      std::memcpy(&rots[9 * (size_t)g], grains[g].rot.data(), 9 * sizeof(double));
      wts[g] = grains[g].weight;
    }
    m.rot = rots.data();
    m.weight = wts.data();
    if (!hgx(hgx_lattice_intensity(gpu_->get(), &m, precision_, d_fc, accumulate ? 1 : 0, d_I), "hgx_lattice_intensity"))
      return false;
    counts_[0] += ng;
    counts_[2] += ng;
    counts_[3] += ng;
    // the call stages the model's host arrays through pinned memory: nothing to wait for here
    return true;
  }

  // any analytic cell: the fused ensemble kernel over all grains, elements and locations
  if (!(flags_ & HGX_HIG_NO_FUSE)) {
    EnsembleBuild b;
    if (ensemble_model(s, grains, b, false, &gs.scaling, &gs.repet)) {
      for (const auto& e : b.el) counts_[2] += (int64_t)ng * e.nloc;
      // the particles sit in the vacuum layer (dnl = 0 above), so q is real and FP32C takes any cell
      if (!hgx(hgx_ensemble_intensity(gpu_->get(), &b.m, precision_, d_fc, accumulate ? 1 : 0, d_I),
               "hgx_ensemble_intensity"))
        return false;
      counts_[0] += ng;
      counts_[3] += ng;
      return true;
    }
  }

  // per grain: form factor (reused while the grain rotation repeats), structure factor, DWBA sum
  double* d_ff = buffer(3, (size_t)nq * 16);
  double* d_eff = buffer(4, (size_t)nq * 16);
  double* d_sf = buffer(5, (size_t)nq * 16);
  if (!d_ff || !d_eff || !d_sf) return fail("device allocation failed");
  for (int g = 0; g < ng; ++g) {
    const HigGrain& G = grains[g];
    bool new_rot = g == 0 || !same9(G.rot.data(), grains[g - 1].rot.data());
    if (new_rot) {
      bool first = true;
      for (const auto& e : uc.elements) {
        const HigShape& sh = in_.shapes.at(e.first);
        const double dn2r = dnl[0] - 2 * sh.delta, dn2i = dnl[1] - 2 * sh.beta;
        // shape rotation rot * Rx * Ry * Rz (src/sim/hipgisaxs_main.cpp:986)
        Rot3 srot = G.rot * Rot3(0, sh.xrot) * Rot3(1, sh.yrot) * Rot3(2, sh.zrot);
        for (const auto& loc : e.second) {
          if (!form_factor(sh, srot, loc, d_eff)) return false;
          if (!hgx(hgx_caxpy(gpu_->get(), nq, dn2r, dn2i, d_eff, first ? 0 : 1, d_ff), "hgx_caxpy")) return false;
          first = false;
        }
      }
    }
    bool new_sf = new_rot || std::memcmp(G.center, grains[g - 1].center, sizeof(G.center)) != 0;
    if (new_sf) {
      ++counts_[3];
      if (!hgx(hgx_sf_lattice(gpu_->get(), la, lb, lc, repet, G.center, 0.0, sf_flags, G.rot.data(), d_sf),
               "hgx_sf_lattice"))
        return false;
    }
    // the reference assigns instead of accumulating in its SAXS branch (:1102-1105), which keeps the
    // last grain only; both experiments accumulate here
    if (!hgx(hgx_dwba_intensity(gpu_->get(), d_fc, d_sf, d_ff, G.weight, (accumulate || g > 0) ? 1 : 0, d_I),
             "hgx_dwba_intensity"))
      return false;
  }
  counts_[1] += ng;
  return true;
}

bool HipGISAXS::run_gisaxs(double alpha_i_deg, double phi_deg, double tilt_deg, std::vector<double>& img) {
  (void)tilt_deg;  // the reference's run_gisaxs reads the tilt angle nowhere
  auto t0 = std::chrono::steady_clock::now();
  const double* d_total = nullptr;
  if (!run_device(alpha_i_deg, phi_deg, &d_total)) return false;
  img.assign((size_t)npix_, 0.0);
  if (d_total && !hgx(hgx_memcpy_d2h(gpu_->get(), img.data(), d_total, (size_t)npix_ * 8), "hgx_memcpy_d2h"))
    return false;
  last_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return true;
}

// The image of one configuration left in device buffer 0; *d_img is null for an input without
// structures (a zero image).
bool HipGISAXS::run_device(double alpha_i_deg, double phi_deg, const double** d_img) {
  *d_img = nullptr;
  if (!loaded_) return fail("no input");
  if (!inited_ && !init()) return false;
  for (auto& c : counts_) c = 0;
  for (auto& c : paths_) c = 0;
  const HigCompute& cp = in_.compute;
  if (cp.nslices > 1) return fail("the slicing scheme (nslices > 1) is not implemented by the reference either");
  // structure correlation (src/sim/hipgisaxs_main.cpp:1205-1262, 1336-1398): nGnE sums iratio-weighted
  // structure intensities; GnE squares each structure's grain sum, GE the iratio-weighted sum. The
  // reference forms those grain sums by reading its summed real grain intensity as if it held one
  // entry per grain, which is defined only for one grain per structure; nGE is not implemented there.
  const std::string& corr = cp.structcorr;
  if (corr == "nGE") return fail("structcorrelation 'nGE' is not implemented by the reference either");
  if (corr != "nGnE" && corr != "GnE" && corr != "GE")
    return fail("unknown structcorrelation '" + corr + "' (nGnE, GnE or GE)");
  for (const auto& s : in_.structures) {
    const HigStructure& st = s.second;
    if (!st.grain_type.empty())
      return fail("structure '" + st.key + "': " + st.grain_type + " structures are not supported");
  }
  const double alphai = alpha_i_deg * PI_ / 180;
  const double phi = phi_deg * PI_ / 180;
  if (in_.structures.empty()) return true;
  // iratio normalisation (src/sim/hipgisaxs_main.cpp:1268-1282)
  std::vector<double> iratio;
  double isum = 0.0;
  for (const auto& s : in_.structures) {
    iratio.push_back(s.second.iratio);
    isum += s.second.iratio;
  }
  if (isum != 1.0)
    for (double& r : iratio) r /= isum;
  double* d_total = buffer(0, (size_t)npix_ * 8);
  double* d_Is = buffer(1, (size_t)npix_ * 8);
  if (!d_total || !d_Is) return fail("device allocation failed");
  const bool prof = (flags_ & HGX_HIG_PROFILE) != 0;
  if (prof && !hgx(hgx_profile_begin(gpu_->get()), "hgx_profile_begin")) return false;
  std::vector<HigGrain> grains;
  size_t si = 0;
  for (const auto& s : in_.structures) {
    if (!make_grains(s.second, rng_, phi, grains)) return false;
    if (corr != "nGnE" && grains.size() != 1)
      return fail("structure '" + s.second.key + "': structcorrelation " + corr +
                  " needs one grain per structure (the reference reads past its grain-intensity array "
                  "for more, src/sim/hipgisaxs_main.cpp:1238-1241, 1254-1257)");
    std::vector<GrainSet> sets;
    if (!grain_sets(s.second, grains, sets)) return false;
    for (size_t k = 0; k < sets.size(); ++k)
      if (!structure_intensity(s.second, alphai, sets[k], k > 0, d_Is)) return false;
    // intensity = sum_s struct_intensity_s * iratio_s (:1351-1364), struct_intensity = I_s^2 for GnE
    if (corr == "GnE") {
      if (!hgx(hgx_axpy_sq(gpu_->get(), npix_, iratio[si], d_Is, si > 0 ? 1 : 0, d_total), "hgx_axpy_sq"))
        return false;
    } else if (!hgx(hgx_axpy(gpu_->get(), npix_, iratio[si], d_Is, si > 0 ? 1 : 0, d_total), "hgx_axpy")) {
      return false;
    }
    ++si;
  }
  // GE: |sum_s iratio_s c_s|^2 with c_s the (real) grain intensity (:1387-1396)
  if (corr == "GE" && !hgx(hgx_axpy_sq(gpu_->get(), npix_, 1.0, d_total, 0, d_total), "hgx_axpy_sq")) return false;
  const double sigma = in_.scattering.smearing;
  if (sigma > DBL_MIN && !(flags_ & HGX_HIG_NO_SMEAR)) {
    int sf = (flags_ & HGX_HIG_SMEAR_COMPAT) ? HGX_SMEAR_COMPAT : 0;
    if (!hgx(hgx_gaussian_smear(gpu_->get(), (int)ncol_, (int)nrow_, sigma, sf, d_total), "hgx_gaussian_smear"))
      return false;
  }
  if (prof && !hgx(hgx_profile_end(gpu_->get(), &kms_[0], &kn_[0], &kms_[1], &kn_[1]), "hgx_profile_end"))
    return false;
  *d_img = d_total;
  return true;
}

void HipGISAXS::last_counts(int64_t counts[4]) const {
  for (int i = 0; i < 4; ++i) counts[i] = counts_[i];
}

bool HipGISAXS::save_gisaxs(const double* data, unsigned ncol, unsigned nrow, const std::string& path) {
  std::ofstream f(path);
  if (!f) return false;
  // the reference streams with the default 6 significant digits; 17 keep every bit
  f << std::setprecision(17);
  for (unsigned z = 0; z < nrow; ++z) {
    for (unsigned y = 0; y < ncol; ++y) f << data[(size_t)ncol * z + y] << "\t";
    f << "\n";
  }
  return (bool)f;
}

bool HipGISAXS::save_tiff(const double* data, unsigned ncol, unsigned nrow, const std::string& path) {
  const size_t n = (size_t)ncol * nrow;
  if (n == 0 || ncol > 65535 || nrow > 65535) return false;
  std::vector<double> v(data, data + n);
  for (double x : v)
    if (!std::isfinite(x)) return false;
  // translate_pixels_to_positive, mat_log10_2d, normalize_pixels (image.cpp:269-312, utilities.cpp:52-75)
  const double mn = *std::min_element(v.begin(), v.end());
  if (mn < 0)
    for (double& x : v) x -= mn;
  for (double& x : v) x = x == 0.0 ? 0.0 : std::log10(x);
  const double lo = *std::min_element(v.begin(), v.end()), hi = *std::max_element(v.begin(), v.end());
  std::vector<uint16_t> px(n);
  for (size_t i = 0; i < n; ++i) {
    const double u = lo == hi ? (lo < 0 ? 0.0 : 1.0) : (v[i] - lo) / (hi - lo);
    px[i] = (uint16_t)std::lround(u * 65535.0);
  }
  // little-endian baseline TIFF: header, pixel strip, one IFD of 10 entries
  std::ofstream f(path, std::ios::binary);
  if (!f) return false;
  auto u16 = [&](uint16_t x) { f.put((char)(x & 0xff)).put((char)(x >> 8)); };
  auto u32 = [&](uint32_t x) {
    for (int k = 0; k < 4; ++k) f.put((char)((x >> (8 * k)) & 0xff));
  };
  const uint32_t strip = 8, bytes = (uint32_t)(n * 2), ifd = strip + bytes + (bytes & 1);
  f.write("II", 2);
  u16(42);
  u32(ifd);
  for (uint16_t x : px) u16(x);
  if (bytes & 1) f.put(0);
  struct Entry {
    uint16_t tag, type;
    uint32_t count, value;
  };
  const Entry e[10] = {{256, 4, 1, ncol},       {257, 4, 1, nrow},  {258, 3, 1, 16}, {259, 3, 1, 1},
                       {262, 3, 1, 1},          {273, 4, 1, strip}, {277, 3, 1, 1},  {278, 4, 1, nrow},
                       {279, 4, 1, bytes},      {284, 3, 1, 1}};
  u16(10);
  for (const Entry& x : e) {
    u16(x.tag);
    u16(x.type);
    u32(x.count);
    if (x.type == 3) {  // SHORT values sit left-justified in the 4-byte field
      u16((uint16_t)x.value);
      u16(0);
    } else {
      u32(x.value);
    }
  }
  u32(0);  // no further IFD
  return (bool)f;
}

}  // namespace hig

// ----- C ABI ---------------------------------------------------------------------------------------

namespace {

int hig_fail(hgx_hig* h, int code) {
  h->err = h->sim.error();
  return code;
}

}  // namespace

extern "C" {

int hgx_hig_create(int device, hgx_hig** out) {
  return hgx_hig_guard(nullptr, [&]() -> int {
    if (!out) return HGX_ERR_INVALID_ARG;
    *out = nullptr;
    hgx_hig* h = new (std::nothrow) hgx_hig(device);
    if (!h) return HGX_ERR_OOM;
    if (!h->sim.error().empty()) {
      delete h;
      return HGX_ERR_DEVICE;
    }
    *out = h;
    return HGX_OK;
  });
}

int hgx_hig_destroy(hgx_hig* h) {
  return hgx_hig_guard(h, [&]() -> int {
    delete h;
    return HGX_OK;
  });
}

int hgx_hig_load(hgx_hig* h, const char* path) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!path) return HGX_ERR_INVALID_ARG;
    return h->sim.construct_input(path) ? HGX_OK : hig_fail(h, HGX_ERR_INVALID_ARG);
  });
}

int hgx_hig_load_text(hgx_hig* h, const char* text, const char* base_dir) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!text) return HGX_ERR_INVALID_ARG;
    return h->sim.construct_input_text(text, base_dir ? base_dir : ".") ? HGX_OK : hig_fail(h, HGX_ERR_INVALID_ARG);
  });
}

int hgx_hig_options(hgx_hig* h, int precision, int flags, unsigned seed) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (precision != HGX_PREC_FP64 && precision != HGX_PREC_FP32C) return HGX_ERR_INVALID_ARG;
    h->sim.options(precision, flags, seed);
    return HGX_OK;
  });
}

int hgx_hig_init(hgx_hig* h) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    return h->sim.init() ? HGX_OK : hig_fail(h, HGX_ERR_INVALID_ARG);
  });
}

int hgx_hig_info(hgx_hig* h, int* ncol, int* nrow, int* nconfig, double* k0) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!h->sim.prepare()) return hig_fail(h, HGX_ERR_INVALID_ARG);
    if (ncol) *ncol = (int)h->sim.ncols();
    if (nrow) *nrow = (int)h->sim.nrows();
    if (nconfig) *nconfig = (int)h->sim.configurations().size();
    if (k0) *k0 = h->sim.k0();
    return HGX_OK;
  });
}

int hgx_hig_config(hgx_hig* h, int config, double* alpha_i, double* phi, double* tilt) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    auto cf = h->sim.configurations();
    if (config < 0 || config >= (int)cf.size()) return HGX_ERR_INVALID_ARG;
    if (alpha_i) *alpha_i = cf[config][0];
    if (phi) *phi = cf[config][1];
    if (tilt) *tilt = cf[config][2];
    return HGX_OK;
  });
}

int hgx_hig_run(hgx_hig* h, int config, double* image) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!image) return HGX_ERR_INVALID_ARG;
    auto cf = h->sim.configurations();
    if (config < 0 || config >= (int)cf.size()) return HGX_ERR_INVALID_ARG;
    std::vector<double> img;
    if (!h->sim.run_gisaxs(cf[config][0], cf[config][1], cf[config][2], img)) return hig_fail(h, HGX_ERR_INVALID_ARG);
    std::memcpy(image, img.data(), img.size() * sizeof(double));
    return HGX_OK;
  });
}

int hgx_hig_compute(hgx_hig* h, double* image) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!image) return HGX_ERR_INVALID_ARG;
    std::vector<double> img;
    if (!h->sim.compute_gisaxs(img)) return hig_fail(h, HGX_ERR_INVALID_ARG);
    std::memcpy(image, img.data(), img.size() * sizeof(double));
    return HGX_OK;
  });
}

int hgx_hig_grains(hgx_hig* h, const char* key, int cap, double* rot, double* center, double* weight, int* n) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!key || !n) return HGX_ERR_INVALID_ARG;
    std::vector<hig::HigGrain> g;
    if (!h->sim.grains(key, g)) return hig_fail(h, HGX_ERR_INVALID_ARG);
    *n = (int)g.size();
    for (int i = 0; i < (int)g.size() && i < cap; ++i) {
      if (rot) std::memcpy(rot + 9 * (size_t)i, g[i].rot.data(), 9 * sizeof(double));
      if (center) std::memcpy(center + 3 * (size_t)i, g[i].center, 3 * sizeof(double));
      if (weight) weight[i] = g[i].weight;
    }
    return HGX_OK;
  });
}

int hgx_hig_last_counts(hgx_hig* h, int64_t counts[4], double* ms) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (counts) h->sim.last_counts(counts);
    if (ms) *ms = h->sim.last_run_ms();
    return HGX_OK;
  });
}

int hgx_hig_last_paths(hgx_hig* h, int64_t paths[3]) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!paths) return HGX_ERR_INVALID_ARG;
    h->sim.last_paths(paths);
    return HGX_OK;
  });
}

int hgx_hig_kernel_ms(hgx_hig* h, double* main_ms, int64_t* main_launches, double* aux_ms, int64_t* aux_launches) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    double m, a;
    int64_t nm, na;
    h->sim.last_kernel_times(&m, &nm, &a, &na);
    if (main_ms) *main_ms = m;
    if (main_launches) *main_launches = nm;
    if (aux_ms) *aux_ms = a;
    if (aux_launches) *aux_launches = na;
    return HGX_OK;
  });
}

int hgx_hig_save(const double* image, int ncol, int nrow, const char* path) {
  return hgx_hig_guard(nullptr, [&]() -> int {
    if (!image || !path || ncol < 1 || nrow < 1) return HGX_ERR_INVALID_ARG;
    return hig::HipGISAXS::save_gisaxs(image, (unsigned)ncol, (unsigned)nrow, path) ? HGX_OK : HGX_ERR_INVALID_ARG;
  });
}

int hgx_hig_save_tiff(const double* image, int ncol, int nrow, const char* path) {
  return hgx_hig_guard(nullptr, [&]() -> int {
    if (!image || !path || ncol < 1 || nrow < 1) return HGX_ERR_INVALID_ARG;
    return hig::HipGISAXS::save_tiff(image, (unsigned)ncol, (unsigned)nrow, path) ? HGX_OK : HGX_ERR_INVALID_ARG;
  });
}

int hgx_hig_read_mesh(const char* path, int* kind, int64_t* ntri, int64_t cap, double* records) {
  return hgx_hig_guard(nullptr, [&]() -> int {
    if (!path || !kind || !ntri) return HGX_ERR_INVALID_ARG;
    std::string p(path), err;
    std::vector<double> rec;
    bool ok;
    if (p.size() > 4 && p.compare(p.size() - 4, 4, ".obj") == 0) {
      *kind = HGX_MESH_EXACT9;
      ok = hig::read_obj_mesh(p, rec, err);
    } else if (p.size() > 4 && p.compare(p.size() - 4, 4, ".dat") == 0) {
      *kind = HGX_MESH_APPROX8;
      ok = hig::read_raw_mesh(p, rec, err);
    } else {
      return HGX_ERR_UNSUPPORTED;
    }
    if (!ok) return HGX_ERR_INVALID_ARG;
    const size_t w = *kind == HGX_MESH_EXACT9 ? 9 : 8;
    *ntri = (int64_t)(rec.size() / w);
    if (records && cap >= *ntri) std::memcpy(records, rec.data(), rec.size() * sizeof(double));
    return HGX_OK;
  });
}

const char* hgx_hig_last_error(const hgx_hig* h) { return h ? h->err.c_str() : "bad handle"; }

}  // extern "C"
