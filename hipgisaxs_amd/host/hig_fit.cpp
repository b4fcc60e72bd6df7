// hig_fit.cpp -- fitting a .hig model to reference data: the parameter-path updates of
// HiGInput::update_params (src/config/hig_input.cpp:2712-2832 and the update_param methods of the
// model classes), the objective of HipGISAXSObjectiveFunction
// (src/analyzer/objective_func_hipgisaxs.cpp:245-332) with its distance computed on the device, and
// the reference's fit drivers: particle swarm (src/analyzer/hipgisaxs_fit_pso.cpp,
// hipgisaxs_fit_pso_particle.cpp), brute force (hipgisaxs_fit_bruteforce.cpp), LMVM
// (hipgisaxs_fit_lmvm.cpp: bounded limited-memory quasi-Newton on central differences) and the plain
// objective evaluation (hipgisaxs_compute_objective.cpp).
#include "hig_fit.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iterator>
#include <limits>
#include <map>
#include <memory>
#include <random>
#include <sstream>

#include "hgx.h"

namespace hig {

namespace {

// extract_first_keyword (src/utils/string_utils.cpp:23-33): text before the first ':' and the rest.
void first_keyword(const std::string& s, std::string& kw, std::string& rest) {
  const size_t pos = s.find(':');
  kw = s.substr(0, pos);
  rest = pos == std::string::npos ? "" : s.substr(pos + 1);
}

// extract_keyword_name_and_key (:65-88): name['key'] -> (name, key); no bracket -> (name, "").
bool name_and_key(const std::string& kw, std::string& name, std::string& key, std::string& err) {
  const size_t pos = kw.find('[');
  if (pos == std::string::npos) {
    name = kw;
    key.clear();
    return true;
  }
  name = kw.substr(0, pos);
  std::string rem = kw.substr(pos + 1);
  if (rem.empty() || rem.back() != ']') {
    err = "ending bracket missing in keyword '" + kw + "'";
    return false;
  }
  rem.pop_back();
  if (rem.size() < 2 || rem.front() != '\'' || rem.back() != '\'') {
    err = "keyword key not enclosed in single-quotes in '" + kw + "'";
    return false;
  }
  key = rem.substr(1, rem.size() - 2);
  return true;
}

// "x"/"0", "y"/"1", "z"/"2" -> 0, 1, 2; else -1
int axis_of(const std::string& s) {
  if (s == "x" || s == "0") return 0;
  if (s == "y" || s == "1") return 1;
  if (s == "z" || s == "2") return 2;
  return -1;
}

}  // namespace

bool HigInput::update_param(const std::string& variable, double v) {
  std::string kw, rest, name, key;
  first_keyword(variable, kw, rest);
  if (kw.empty()) {
    err_ = "empty parameter variable";
    return false;
  }
  if (!name_and_key(kw, name, key, err_)) return false;
  const std::string where = "in parameter variable '" + variable + "'";
  auto bad = [&](const std::string& what) {
    err_ = "invalid keyword '" + what + "' " + where;
    return false;
  };
  std::string k1, r1, n1, key1;
  first_keyword(rest, k1, r1);
  if (!name_and_key(k1, n1, key1, err_)) return false;
  std::string k2, r2;
  first_keyword(r1, k2, r2);

  if (name == "shape") {  // Shape::update_param (src/model/shape.cpp:190-247)
    auto it = shapes.find(key);
    if (it == shapes.end()) return bad(kw);
    HigShape& sh = it->second;
    if (n1 == "originvec") {
      const int ax = axis_of(k2);
      if (ax < 0) return bad(k2);
      sh.origin[ax] = v;
    } else if (n1 == "xrot" || n1 == "yrot" || n1 == "zrot") {
      // the value is stored as given (radians), as Shape::update_param stores it; the file's angles
      // are degrees. (The reference's xrot case falls through into the param case.)
      (n1 == "xrot" ? sh.xrot : n1 == "yrot" ? sh.yrot : sh.zrot) = v;
    } else if (n1 == "param") {  // ShapeParam::update_param (:145-187)
      auto pit = sh.params.find(key1);
      if (pit == sh.params.end()) return bad(k1);
      HigShapeParam& p = pit->second;
      if (k2 == "min") p.min = v;
      else if (k2 == "max") p.max = v;
      else if (k2 == "p1") p.p1 = v;
      else if (k2 == "p2") p.p2 = v;
      else if (k2 == "nvalues") p.nvalues = (int)v;
      else if (k2 == "stat") return true;  // immutable, ignored
      else return bad(k2);
    } else {
      return bad(k1);
    }
    return true;
  }
  if (name == "layer") {  // Layer::update_param (src/model/layer.cpp:41-86)
    HigLayer* L = nullptr;
    for (auto& l : layers)
      if (l.second.key == key) L = &l.second;
    if (!L) return bad(kw);
    if (n1 == "thickness") L->thickness = v;
    else if (n1 == "refindex" && k2 == "delta") L->delta = v;
    else if (n1 == "refindex" && k2 == "beta") L->beta = v;
    else if (n1 == "key" || n1 == "order") return true;  // immutable, ignored
    else return bad(n1 == "refindex" ? k2 : k1);
    return true;
  }
  if (name == "structure") {  // Structure::update_param (src/model/structure.cpp:519-820)
    auto it = structures.find(key);
    if (it == structures.end()) return bad(kw);
    HigStructure& st = it->second;
    std::string k3, r3;
    first_keyword(r2, k3, r3);
    if (n1 == "iratio") {
      st.iratio = v;
    } else if (n1 == "key") {
      return true;
    } else if (n1 == "grain") {
      if (k2 == "refindex") {
        // the grain's refractive index is stored by the reference and read by no simulation step
        if (k3 != "delta" && k3 != "beta") return bad(k3);
      } else if (k2 == "lattice") {
        if (k3 == "a" || k3 == "b" || k3 == "c") {
          std::string k4, r4;
          first_keyword(r3, k4, r4);
          const int ax = axis_of(k4);
          if (ax < 0) return bad(k4);
          if (!st.abc_set) {  // the lattice type's vectors, then the component (Lattice::ax ...)
            std::string e;
            if (!lattice_vectors(st, st.a, st.b, st.c, e)) {
              err_ = e;
              return false;
            }
            st.abc_set = true;
          }
          (k3 == "a" ? st.a : k3 == "b" ? st.b : st.c)[ax] = v;
        } else if (k3 == "type" || k3 == "hkl" || k3 == "abangle" || k3 == "caratio") {
          return true;  // immutable, or read by no lattice construction of the reference
        } else {
          return bad(k3);
        }
      } else if (k2 == "scaling") {
        if (k3.empty()) {
          st.scaling = {v, v, v};
        } else {
          const int ax = k3 == "a" ? 0 : k3 == "b" ? 1 : k3 == "c" ? 2 : -1;
          if (ax < 0) return bad(k3);
          std::string k4, r4;
          first_keyword(r3, k4, r4);
          if (k4 == "mean") st.scaling[ax] = v;
          else if (k4 == "stddev" || k4 == "nsamples") {
            err_ = "scaling distributions are not supported (" + variable + ")";
            return false;
          } else {
            return bad(k4);
          }
        }
      } else if (k2 == "transvec" || k2 == "repetition") {
        const int ax = axis_of(k3);
        if (ax < 0) return bad(k3);
        (k2 == "transvec" ? st.transvec : st.repetition)[ax] = v;
      } else {
        return bad(k2);
      }
    } else if (n1 == "ensemble") {
      if (k2 == "spacing" || k2 == "maxgrains") {
        const int ax = axis_of(k3);
        if (ax < 0) return bad(k3);
        if (k2 == "maxgrains" && !(v >= 1 && v <= 1e6 && v == std::floor(v))) {
          err_ = "maxgrains must be a positive integer up to 1e6 (" + variable + ")";
          return false;
        }
        (k2 == "spacing" ? st.spacing : st.maxgrains)[ax] = v;
      } else if (k2 == "orientations" || k2 == "distribution") {
        return true;  // ignored by the reference
      } else {
        return bad(k2);
      }
    } else {
      return bad(k1);
    }
    return true;
  }
  if (name == "instrumentation") {
    if (n1 == "detector") return true;  // every detector value is immutable (inst_detector.cpp:38-65)
    if (n1 != "scattering") return bad(k1);
    std::string k3, r3;
    first_keyword(r2, k3, r3);
    HigScattering& sc = scattering;  // ScatteringParams::update_param (inst_scattering.cpp:41-167)
    if (k2 == "alphai" || k2 == "inplanerot" || k2 == "tilt") {
      const int a = k2 == "alphai" ? 0 : k2 == "inplanerot" ? 1 : 2;
      double* mn[3] = {&sc.alphai_min, &sc.phi_min, &sc.tilt_min};
      double* mx[3] = {&sc.alphai_max, &sc.phi_max, &sc.tilt_max};
      double* st[3] = {&sc.alphai_step, &sc.phi_step, &sc.tilt_step};
      if (k3 == "min") *mn[a] = v;
      else if (k3 == "max") *mx[a] = v;
      else if (k3 == "step") *st[a] = v;
      else return bad(k3);
    } else if (k2 == "photon") {
      if (k3 == "value") sc.photon = v;
      else if (k3 != "unit") return bad(k3);
    } else if (k2 == "expt" || k2 == "polarization" || k2 == "smearing" || k2 == "coherence" ||
               k2 == "spotarea") {
      return true;  // immutable, or read by no simulation step
    } else {
      return bad(k2);
    }
    return true;
  }
  if (name == "computation") return true;  // ComputeParams::update_param: every value immutable
  return bad(name);
}

// ----- the objective ----------------------------------------------------------------------------------

namespace {

// ImageData::read (src/analyzer/ImageData.cpp:79-104): one image row per line, the column count from
// the first line.
bool read_ascii_image(const std::string& path, std::vector<double>& data, unsigned& ncol, unsigned& nrow,
                      std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "unable to open reference data file " + path;
    return false;
  }
  data.clear();
  ncol = nrow = 0;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ls(line);
    double x;
    size_t before = data.size();
    while (ls >> x) data.push_back(x);
    const size_t got = data.size() - before;
    if (got == 0) continue;
    if (nrow == 0) ncol = (unsigned)got;
    else if (got != ncol) {
      err = "reference data row " + std::to_string(nrow) + " has " + std::to_string(got) + " values, not " +
            std::to_string(ncol);
      return false;
    }
    ++nrow;
  }
  if (nrow == 0) {
    err = "reference data file " + path + " holds no values";
    return false;
  }
  return true;
}

// EDFReader (src/file/edf_reader.cpp:24-150): a header of `key = value ;` items (white space dropped)
// closed by '}', read in 512-byte blocks; the data start at the block after the one holding the '}'
// and are EDF_BinarySize bytes of 32-bit floats in the machine's (little-endian) order, Dim_1 columns
// by Dim_2 rows, whatever DataType and ByteOrder say.
bool read_edf_image(const std::string& path, std::vector<double>& data, unsigned& ncol, unsigned& nrow,
                    std::string& err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "could not open the EDF file " + path;
    return false;
  }
  const std::string bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  auto white = [](char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r'; };
  std::map<std::string, std::string> header;
  size_t p = 0, close = std::string::npos;
  while (close == std::string::npos) {
    std::string key, value;
    bool got_key = false;
    for (; p < bytes.size(); ++p) {  // the key, up to '='; a '}' here ends the header
      const char c = bytes[p];
      if (c == '=') {
        got_key = true;
        ++p;
        break;
      }
      if (c == '}') {
        close = p;
        break;
      }
      if (!white(c)) key += c;
    }
    if (close != std::string::npos) break;
    if (!got_key) {
      err = "EDF file " + path + ": the header is not closed";
      return false;
    }
    bool got_value = false;
    for (; p < bytes.size(); ++p) {
      const char c = bytes[p];
      if (c == ';') {
        got_value = true;
        ++p;
        break;
      }
      if (!white(c)) value += c;
    }
    if (!got_value) {
      err = "EDF file " + path + ": a header item has no ';'";
      return false;
    }
    header[key] = value;
    size_t q = p;  // header_done: the next character that is not white space closes it
    while (q < bytes.size() && white(bytes[q])) ++q;
    if (q < bytes.size() && bytes[q] == '}') close = q;
  }
  auto num = [&](const char* k) { return std::atol(header[k].c_str()); };  // atoi / atol as the reference
  const long cols = num("Dim_1"), rows = num("Dim_2"), nbytes = num("EDF_BinarySize");
  const size_t start = (close / 512 + 1) * 512;
  if (cols <= 0 || rows <= 0 || nbytes < 0 || (size_t)nbytes / 4 != (size_t)cols * (size_t)rows) {
    err = "EDF file " + path + ": mismatch in reference data size (Dim_1 x Dim_2 against EDF_BinarySize / 4)";
    return false;
  }
  if (start + (size_t)nbytes > bytes.size()) {
    err = "EDF file " + path + ": shorter than its header says";
    return false;
  }
  data.resize((size_t)cols * rows);
  for (size_t i = 0; i < data.size(); ++i) {
    float v;
    std::memcpy(&v, bytes.data() + start + 4 * i, 4);
    data[i] = v;
  }
  ncol = (unsigned)cols;
  nrow = (unsigned)rows;
  return true;
}

}  // namespace

bool HipGISAXS::update_param(const std::string& variable, double value) {
  if (!loaded_) return fail("no input");
  if (!in_.update_param(variable, value)) return fail(in_.error());
  if (variable.compare(0, 15, "instrumentation") == 0) {
    inited_ = false;  // k0 and the q-grid follow the new instrument values at the next run
    grid_ok_ = false;
  }
  return true;
}

std::vector<std::string> HipGISAXS::fit_keys() const {
  std::vector<std::string> k;
  for (const auto& p : in_.fit_params) k.push_back(p.first);
  return k;
}

// HipGISAXSObjectiveFunction::set_reference_data (objective_func_hipgisaxs.cpp:110-170) for the
// first reference data set.
bool HipGISAXS::fit_init() {
  fit_ready_ = false;
  history_.clear();
  if (!loaded_) return fail("no input");
  if (!in_.reference_set || in_.reference.path.empty()) return fail("the input has no referencedata path");
  const HigFitReference& ref = in_.reference;
  const size_t dot = ref.path.find_last_of('.');
  const std::string ext = dot == std::string::npos ? "" : ref.path.substr(dot + 1);
  const bool edf = ext == "edf" || ext == "EDF";
  if (!edf && ext != "txt" && ext != "TXT" && ext != "dat" && ext != "DAT" && ext != "out" && ext != "OUT")
    return fail("error in reference data file format (extension of '" + ref.path + "')");
  std::vector<double> O;
  unsigned ncol = 0, nrow = 0;
  std::string err;
  if (!(edf ? read_edf_image(ref.path, O, ncol, nrow, err) : read_ascii_image(ref.path, O, ncol, nrow, err)))
    return fail(err);
  if (!prepare()) return false;
  if (ncol != ncol_ || nrow != nrow_) {
    // override_qregion (src/sim/hipgisaxs_main.cpp:236-276): the reference's region and size
    if (ref.region != "qspace") return fail("override of a '" + ref.region + "' region is not implemented");
    in_.compute.minpoint[0] = ref.regmin[0];
    in_.compute.minpoint[1] = ref.regmin[1];
    in_.compute.maxpoint[0] = ref.regmax[0];
    in_.compute.maxpoint[1] = ref.regmax[1];
    in_.compute.resolution[0] = (int)ncol;
    in_.compute.resolution[1] = (int)nrow;
    inited_ = false;
    grid_ok_ = false;
    if (!prepare()) return false;
  }
  std::vector<double> mask((size_t)npix_, 1.0);
  if (!ref.mask.empty()) {  // read_mask_data (:174-193)
    std::ifstream f(ref.mask);
    if (!f) return fail("could not open mask data file " + ref.mask);
    std::vector<double> m;
    double x;
    while (f >> x) m.push_back(x);
    if ((int64_t)m.size() != npix_)
      return fail("mask and reference data dimension sizes do not match [" + std::to_string(ncol_) + " * " +
                  std::to_string(nrow_) + "] != " + std::to_string(m.size()));
    mask = m;
  }
  if (!device()) return false;
  double* d_O = buffer(6, (size_t)npix_ * 8);
  double* d_m = buffer(7, (size_t)npix_ * 8);
  if (!d_O || !d_m) return fail("device allocation failed");
  if (!hgx(hgx_memcpy_h2d(gpu_->get(), d_O, O.data(), (size_t)npix_ * 8), "hgx_memcpy_h2d") ||
      !hgx(hgx_memcpy_h2d(gpu_->get(), d_m, mask.data(), (size_t)npix_ * 8), "hgx_memcpy_h2d"))
    return false;
  fit_npix_ = npix_;
  fit_ready_ = true;
  return true;
}

bool HipGISAXS::fit_objective(const double* x, int metric, double reg, double* value) {
  if (!fit_ready_) return fail("fit_init has not run");
  if (configurations().size() != 1)
    return fail("currently you can simulate only for single alpha_i, phi and tilt angles for fitting");
  std::vector<double> xv;
  double pnorm = 0.0;
  size_t i = 0;
  for (const auto& p : in_.fit_params) {
    if (!update_param(p.second.variable, x[i])) return false;
    // regularization about the range midpoint (param_space_mean, include/config/hig_input.hpp:322-325)
    const double mean = (p.second.max - p.second.min) / 2.0 + p.second.min;
    pnorm += (x[i] - mean) * (x[i] - mean);
    xv.push_back(x[i]);
    ++i;
  }
  if (!inited_ && !init()) return false;
  if (npix_ != fit_npix_) return fail("the image size changed after fit_init");
  const HigScattering& sc = in_.scattering;
  const double* d_img = nullptr;
  if (!run_device(sc.alphai_min, sc.phi_min, &d_img)) return false;
  double D = 0.0;
  if (!d_img) {  // no structures: a zero image
    double* d_zero = buffer(0, (size_t)npix_ * 8);
    const std::vector<double> zeros((size_t)npix_, 0.0);
    if (!d_zero || !hgx(hgx_memcpy_h2d(gpu_->get(), d_zero, zeros.data(), (size_t)npix_ * 8), "hgx_memcpy_h2d"))
      return false;
    d_img = d_zero;
  }
  if (!hgx(hgx_image_distance(gpu_->get(), npix_, metric, d_img, static_cast<const double*>(buf_[6]),
                              static_cast<const double*>(buf_[7]), &D),
           "hgx_image_distance"))
    return false;
  // the regularization spread over the (one-element) distance vector; non-finite -> 1e8 (:297-312)
  D = std::isfinite(D) ? D + (reg / 2) * pnorm : 1e8;
  *value = D;
  xv.push_back(D);
  history_.push_back(xv);
  return true;
}

namespace {

// Whether two builds of every structure have the same counts (so that their difference is a tangent).
bool same_layout(const std::vector<std::unique_ptr<EnsembleBuild>>& A, const std::vector<std::unique_ptr<EnsembleBuild>>& B) {
  if (A.size() != B.size()) return false;
  for (size_t s = 0; s < A.size(); ++s) {
    const hgx_ensemble_model &a = A[s]->m, &b = B[s]->m;
    if (a.nelem != b.nelem || a.ngrains != b.ngrains || a.sf_flags != b.sf_flags) return false;
    for (int k = 0; k < 3; ++k)
      if (a.repet[k] != b.repet[k]) return false;
    for (int e = 0; e < a.nelem; ++e) {
      const hgx_ensemble_element &x = a.elem[e], &y = b.elem[e];
      if (x.shape != y.shape || x.nloc != y.nloc || x.params.flags != y.params.flags) return false;
      for (int d = 0; d < 4; ++d)
        if (x.params.n[d] != y.params.n[d]) return false;
    }
  }
  return true;
}

// (P - M) / w field by field into D (same layout).
void tangent_build(const EnsembleBuild& P, const EnsembleBuild& M, double w, EnsembleBuild& D) {
  auto diff = [w](const std::vector<double>& a, const std::vector<double>& b) {
    std::vector<double> r(a.size());
    for (size_t i = 0; i < a.size(); ++i) r[i] = (a[i] - b[i]) / w;
    return r;
  };
  D.el = P.el;
  D.stores.resize(P.stores.size());
  for (size_t i = 0; i < P.stores.size(); ++i) D.stores[i] = diff(P.stores[i], M.stores[i]);
  D.locs.resize(P.locs.size());
  for (size_t i = 0; i < P.locs.size(); ++i) D.locs[i] = diff(P.locs[i], M.locs[i]);
  for (size_t e = 0; e < P.el.size(); ++e) {
    hgx_ensemble_element& t = D.el[e];
    const hgx_ensemble_element &a = P.el[e], &b = M.el[e];
    t.params.tau = (a.params.tau - b.params.tau) / w;
    t.params.eta = (a.params.eta - b.params.eta) / w;
    for (int k = 0; k < 9; ++k) t.rot[k] = (a.rot[k] - b.rot[k]) / w;
    t.dn2_re = (a.dn2_re - b.dn2_re) / w;
    t.dn2_im = (a.dn2_im - b.dn2_im) / w;
  }
  D.rots = diff(P.rots, M.rots);
  D.cen = diff(P.cen, M.cen);
  D.wts = diff(P.wts, M.wts);
  D.m = P.m;
  for (int k = 0; k < 3; ++k) {
    D.m.la[k] = (P.m.la[k] - M.m.la[k]) / w;
    D.m.lb[k] = (P.m.lb[k] - M.m.lb[k]) / w;
    D.m.lc[k] = (P.m.lc[k] - M.m.lc[k]) / w;
  }
  D.m.dw_sigma = (P.m.dw_sigma - M.m.dw_sigma) / w;
  D.relink();
}

}  // namespace

bool HipGISAXS::fit_gradient(const double* x, int metric, double reg, double* value, double* grad, int* computed) {
  if (!fit_ready_) return fail("fit_init has not run");
  if (configurations().size() != 1)
    return fail("currently you can simulate only for single alpha_i, phi and tilt angles for fitting");
  const size_t np = in_.fit_params.size();
  std::vector<std::string> vars;
  std::vector<double> mean, lo, hi;
  double pnorm = 0.0;
  size_t i = 0;
  for (const auto& p : in_.fit_params) {
    if (!update_param(p.second.variable, x[i])) return false;
    vars.push_back(p.second.variable);
    mean.push_back((p.second.max - p.second.min) / 2.0 + p.second.min);
    lo.push_back(p.second.min);
    hi.push_back(p.second.max);
    pnorm += (x[i] - mean[i]) * (x[i] - mean[i]);
    ++i;
  }
  for (size_t j = 0; j < np; ++j) {
    computed[j] = 0;
    grad[j] = 0.0;
  }
  if (!inited_ && !init()) return false;
  if (npix_ != fit_npix_) return fail("the image size changed after fit_init");
  // what the tangent pass covers: structures of analytic cells, the run_device checks passed
  // (truncated spheres have no tangent: those inputs take central differences)
  bool ok = !in_.structures.empty() && in_.compute.nslices <= 1 && in_.compute.structcorr == "nGnE" &&
            !(flags_ & HGX_HIG_TRUNCATED_SPHERE);
  for (const auto& s : in_.structures)
    ok = ok && s.second.grain_type.empty() && !s.second.scaling_dist && !s.second.repetition_dist;
  const double kDeg = 3.141592653589793 / 180;
  const HigScattering& sc = in_.scattering;
  const double alphai = sc.alphai_min * kDeg, phi = sc.phi_min * kDeg;
  const GlibcRandom rng0 = rng_;
  // every structure's model at the current values, grains drawn from rng0 as one run would draw them
  auto build_all = [&](std::vector<std::unique_ptr<EnsembleBuild>>& out, std::vector<double>& ir,
                       GlibcRandom& rng) -> bool {
    out.clear();
    for (const auto& s : in_.structures) {
      std::vector<HigGrain> grains;
      if (!make_grains(s.second, rng, phi, grains)) return false;
      auto b = std::make_unique<EnsembleBuild>();
      if (!ensemble_model(s.second, grains, *b, false)) return false;
      out.push_back(std::move(b));
    }
    ir = iratios();
    return true;
  };
  std::vector<std::unique_ptr<EnsembleBuild>> B0;
  std::vector<double> ir0;
  GlibcRandom rng = rng0;
  ok = ok && build_all(B0, ir0, rng);
  const GlibcRandom rng_after = rng;
  // directions: variables that move the model's tables only, with builds of equal layout at x +- h
  std::vector<int> dirs;
  std::vector<std::vector<std::unique_ptr<EnsembleBuild>>> T;
  std::vector<std::vector<double>> Tir;
  for (size_t j = 0; ok && j < np; ++j) {
    const std::string& v = vars[j];
    if (v.compare(0, 15, "instrumentation") == 0 || v.compare(0, 5, "layer") == 0) continue;
    // central differences of the host tables: they are smooth (mostly linear) in the variable, so a
    // small step leaves only rounding, about 1e-16 |table| / h
    const double h = 1e-6 * std::max({std::fabs(x[j]), 0.01 * std::fabs(hi[j] - lo[j]), 1e-3});
    std::vector<std::unique_ptr<EnsembleBuild>> Bp, Bm;
    std::vector<double> irp, irm;
    GlibcRandom rp = rng0, rm = rng0;
    const bool okp = update_param(v, x[j] + h) && build_all(Bp, irp, rp);
    const bool okm = update_param(v, x[j] - h) && build_all(Bm, irm, rm);
    if (!update_param(v, x[j])) return false;
    err_.clear();
    if (!okp || !okm || !same_layout(B0, Bp) || !same_layout(B0, Bm)) continue;
    std::vector<std::unique_ptr<EnsembleBuild>> D;
    for (size_t s = 0; s < B0.size(); ++s) {
      D.push_back(std::make_unique<EnsembleBuild>());
      tangent_build(*Bp[s], *Bm[s], 2 * h, *D.back());
    }
    std::vector<double> dir(ir0.size());
    for (size_t s = 0; s < ir0.size(); ++s) dir[s] = (irp[s] - irm[s]) / (2 * h);
    dirs.push_back((int)j);
    T.push_back(std::move(D));
    Tir.push_back(dir);
  }
  err_.clear();
  rng_ = rng0;
  if (dirs.empty()) return fit_objective(x, metric, reg, value);  // the caller differentiates
  const int nd = (int)dirs.size();
  const bool gisaxs = sc.expt == "gisaxs";
  if (!ensure_grid(gisaxs ? alphai : -1.0, 0.0, 0.0)) return false;
  double* d_fc = nullptr;
  if (!fresnel_coeffs(alphai, &d_fc)) return false;
  double* d_total = buffer(0, (size_t)npix_ * 8);
  double* d_Is = buffer(1, (size_t)npix_ * 8);
  double* d_dtot = buffer(3, (size_t)nd * npix_ * 8);
  double* d_dIs = buffer(4, (size_t)nd * npix_ * 8);
  if (!d_total || !d_Is || !d_dtot || !d_dIs) return fail("device allocation failed");
  hgx_ctx* ctx = gpu_->get();
  std::vector<hgx_ensemble_model> tm(nd);
  for (size_t s = 0; s < B0.size(); ++s) {
    for (int k = 0; k < nd; ++k) tm[k] = T[k][s]->m;
    if (!hgx(hgx_ensemble_intensity_tangent(ctx, &B0[s]->m, nd, tm.data(), d_fc, d_Is, d_dIs),
             "hgx_ensemble_intensity_tangent"))
      return false;
    // intensity = sum_s iratio_s I_s (src/sim/hipgisaxs_main.cpp:1351-1364) and its tangents
    if (!hgx(hgx_axpy(ctx, npix_, ir0[s], d_Is, s > 0 ? 1 : 0, d_total), "hgx_axpy") ||
        !hgx(hgx_axpy(ctx, (int64_t)nd * npix_, ir0[s], d_dIs, s > 0 ? 1 : 0, d_dtot), "hgx_axpy"))
      return false;
    for (int k = 0; k < nd; ++k)
      if (Tir[k][s] != 0.0 && !hgx(hgx_axpy(ctx, npix_, Tir[k][s], d_Is, 1, d_dtot + (size_t)k * npix_), "hgx_axpy"))
        return false;
  }
  // the smearing is linear: the same operator on the image and on every tangent
  if (sc.smearing > DBL_MIN && !(flags_ & HGX_HIG_NO_SMEAR)) {
    const int sf = (flags_ & HGX_HIG_SMEAR_COMPAT) ? HGX_SMEAR_COMPAT : 0;
    for (int k = -1; k < nd; ++k)
      if (!hgx(hgx_gaussian_smear(ctx, (int)ncol_, (int)nrow_, sc.smearing, sf, k < 0 ? d_total : d_dtot + (size_t)k * npix_),
               "hgx_gaussian_smear"))
        return false;
  }
  double D = 0.0;
  std::vector<double> dD(nd);
  if (!hgx(hgx_image_distance_tangent(ctx, npix_, metric, d_total, static_cast<const double*>(buf_[6]),
                                      static_cast<const double*>(buf_[7]), nd, d_dtot, &D, dD.data()),
           "hgx_image_distance_tangent"))
    return false;
  rng_ = rng_after;
  const bool finite = std::isfinite(D);
  double f = finite ? D + (reg / 2) * pnorm : 1e8;
  for (int k = 0; k < nd; ++k) {
    const int j = dirs[k];
    grad[j] = finite ? dD[k] + reg * (x[j] - mean[j]) : 0.0;
    computed[j] = 1;
  }
  if (precision_ == HGX_PREC_FP32C) {
    // the value in the working precision, so that it compares with the fit's other evaluations
    rng_ = rng0;
    return fit_objective(x, metric, reg, value);
  }
  *value = f;
  std::vector<double> row(x, x + np);
  row.push_back(f);
  history_.push_back(row);
  return true;
}

bool HipGISAXS::fit_objective_many(const std::vector<std::vector<double>>& X, int metric, double reg,
                                   std::vector<double>& out) {
  out.clear();
  if (!fit_ready_) return fail("fit_init has not run");
  if (configurations().size() != 1)
    return fail("currently you can simulate only for single alpha_i, phi and tilt angles for fitting");
  const size_t n = X.size(), np = in_.fit_params.size();
  if (n == 0) return true;
  if (!device()) return false;
  double* d_D = buffer(8, n * sizeof(double));
  if (!d_D) return fail("device allocation failed");
  std::vector<double> pnorm(n, 0.0);
  for (size_t r = 0; r < n; ++r) {
    if (X[r].size() != np) return fail("fit_objective_many: a row has the wrong length");
    size_t i = 0;
    for (const auto& p : in_.fit_params) {
      if (!update_param(p.second.variable, X[r][i])) return false;
      const double mean = (p.second.max - p.second.min) / 2.0 + p.second.min;
      pnorm[r] += (X[r][i] - mean) * (X[r][i] - mean);
      ++i;
    }
    if (!inited_ && !init()) return false;
    if (npix_ != fit_npix_) return fail("the image size changed after fit_init");
    const HigScattering& sc = in_.scattering;
    const double* d_img = nullptr;
    if (!run_device(sc.alphai_min, sc.phi_min, &d_img)) return false;
    if (!d_img) {  // no structures: a zero image
      double* d_zero = buffer(0, (size_t)npix_ * 8);
      const std::vector<double> zeros((size_t)npix_, 0.0);
      if (!d_zero || !hgx(hgx_memcpy_h2d(gpu_->get(), d_zero, zeros.data(), (size_t)npix_ * 8), "hgx_memcpy_h2d"))
        return false;
      d_img = d_zero;
    }
    // the next row's image overwrites buffer 0 only after this distance has read it (one stream)
    if (!hgx(hgx_image_distance_async(gpu_->get(), npix_, metric, d_img, static_cast<const double*>(buf_[6]),
                                      static_cast<const double*>(buf_[7]), d_D + r),
             "hgx_image_distance_async"))
      return false;
  }
  std::vector<double> D(n);
  if (!hgx(hgx_memcpy_d2h(gpu_->get(), D.data(), d_D, n * sizeof(double)), "hgx_memcpy_d2h")) return false;
  out.resize(n);
  for (size_t r = 0; r < n; ++r) {
    out[r] = std::isfinite(D[r]) ? D[r] + (reg / 2) * pnorm[r] : 1e8;
    std::vector<double> row = X[r];
    row.push_back(out[r]);
    history_.push_back(row);
  }
  return true;
}

// ----- fit drivers -----------------------------------------------------------------------------------

int metric_of(const std::string& name) {
  if (name == "sqrt_unit_l2") return HGX_METRIC_SQRT_UNIT_L2;
  if (name == "sqrt_c_l2") return HGX_METRIC_SQRT_C_L2;
  if (name == "cbrt_unit_l2") return HGX_METRIC_CBRT_UNIT_L2;
  if (name == "cbrt_c_l2") return HGX_METRIC_CBRT_C_L2;
  if (name == "chi2") return HGX_METRIC_CHI2;
  if (name == "l2") return HGX_METRIC_L2;
  return -1;
}

namespace {

struct Problem {
  HipGISAXS& sim;
  int metric;
  double reg;
  std::vector<double> lo, hi, step, x0;
  int64_t evals = 0;
  bool fd_only = false;     // LMVM gradients by central differences only (the reference's way)
  int64_t grad_passes = 0;  // tangent passes of LMVM
  bool eval(const std::vector<double>& x, double& f) {
    ++evals;
    return sim.fit_objective(x.data(), metric, reg, &f);
  }
  bool eval_many(const std::vector<std::vector<double>>& X, std::vector<double>& f) {
    evals += (int64_t)X.size();
    return sim.fit_objective_many(X, metric, reg, f);
  }
};

// ParticleSwarmOptimization, base (inertia weight) flavour: particles drawn uniformly in the
// limits with velocities uniform in [-max, max] (PSOParticle::init_random_uniform), each generation
// evaluated particle by particle in index order, personal and global bests updated after each
// evaluation, then every particle moves (update_particle: v = omega v + phi1 r1 (pbest - x) +
// phi2 r2 (gbest - x), v clamped to [-max, max], x clamped to [min, max]); omega halves per
// generation when tuned (src/analyzer/hipgisaxs_fit_pso.cpp:123-211, 425-660).
bool run_pso(Problem& P, const HigFitAlgorithm& a, uint64_t seed, std::vector<double>& best, double& fbest,
             std::string& err) {
  auto need = [&](const char* k, double& out) {
    auto it = a.params.find(k);
    if (it == a.params.end()) {
      err = std::string("mandatory PSO parameter '") + k + "' missing";
      return false;
    }
    out = it->second;
    return true;
  };
  double omega, phi1, phi2, np_d, ng_d;
  if (!need("pso_omega", omega) || !need("pso_phi1", phi1) || !need("pso_phi2", phi2) ||
      !need("pso_num_particles", np_d) || !need("pso_num_generations", ng_d))
    return false;
  auto it = a.params.find("pso_type");
  if (it != a.params.end() && it->second != 0.0) {
    err = "only the base PSO flavour (pso_type 0) is implemented";
    return false;
  }
  const bool tune = a.params.count("pso_tune_omega") && a.params.at("pso_tune_omega") > 0.5;
  const int np = (int)np_d, ng = (int)ng_d, d = (int)P.lo.size();
  if (np < 1 || ng < 0 || np > 100000) {
    err = "pso_num_particles must be in [1, 100000]";
    return false;
  }
  std::vector<double> xmin(d), xmax(d);
  for (int j = 0; j < d; ++j) {
    xmin[j] = std::min(P.lo[j], P.hi[j]);
    xmax[j] = std::max(P.lo[j], P.hi[j]);
  }
  MtUniform rng(seed);
  std::vector<std::vector<double>> x(np, std::vector<double>(d)), v = x, pb = x;
  std::vector<double> pf(np, std::numeric_limits<double>::max());
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < d; ++j) {
      x[i][j] = xmin[j] + rng.next() * (xmax[j] - xmin[j]);
      v[i][j] = -xmax[j] + rng.next() * (xmax[j] + xmax[j]);
      pb[i][j] = x[i][j];
    }
  best.assign(d, 0.0);
  fbest = std::numeric_limits<double>::max();
  const bool sequential = (P.sim.flags() & HGX_HIG_FIT_SEQUENTIAL) != 0;
  std::vector<double> fgen;
  for (int gen = 0; gen < ng; ++gen) {
    // the generation's objectives (one synchronisation), then the best-so-far updates in particle
    // order, as the reference's loop makes them
    if (!sequential && !P.eval_many(x, fgen)) return false;
    for (int i = 0; i < np; ++i) {
      double f;
      if (sequential) {
        if (!P.eval(x[i], f)) return false;
      } else {
        f = fgen[i];
      }
      if (pf[i] > f) {
        pf[i] = f;
        pb[i] = x[i];
      }
      if (fbest > f) {
        fbest = f;
        best = x[i];
      }
    }
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < d; ++j) {
        const double r1 = rng.next(), r2 = rng.next();
        const double nv = omega * v[i][j] + phi1 * r1 * (pb[i][j] - x[i][j]) + phi2 * r2 * (best[j] - x[i][j]);
        v[i][j] = std::min(std::max(nv, -xmax[j]), xmax[j]);
        x[i][j] = std::min(std::max(x[i][j] + v[i][j], xmin[j]), xmax[j]);
      }
    if (tune) omega /= 2.0;
  }
  return true;
}

// BruteForceOptimization::loop_over_param (hipgisaxs_fit_bruteforce.cpp:59-72): every point of the
// grid p = min, min + step, ... <= max (the step accumulated; a negative step means 1), first
// parameter outermost.
bool run_bruteforce(Problem& P, std::vector<double>& best, double& fbest, std::string& err) {
  const int d = (int)P.lo.size();
  double count = 1.0;
  for (int j = 0; j < d; ++j) {
    const double st = P.step[j] < 0 ? 1.0 : P.step[j];
    if (!(P.lo[j] <= P.hi[j])) {  // the grid would hold no point
      err = "brute force needs min <= max for every fit parameter";
      return false;
    }
    if (!(st > 0)) {
      err = "brute force needs positive steps";
      return false;
    }
    count *= std::max(1.0, std::floor((P.hi[j] - P.lo[j]) / st) + 1.0);
  }
  if (count > 1e7) {
    err = "brute-force grid of " + std::to_string(count) + " points is too large";
    return false;
  }
  fbest = std::numeric_limits<double>::max();
  std::vector<double> cur;
  std::function<bool(int)> loop = [&](int j) -> bool {
    if (j == d) {
      double f;
      if (!P.eval(cur, f)) return false;
      if (f < fbest) {
        fbest = f;
        best = cur;
      }
      return true;
    }
    const double st = P.step[j] < 0 ? 1.0 : P.step[j];
    for (double p = P.lo[j]; p <= P.hi[j]; p += st) {
      cur.push_back(p);
      if (!loop(j + 1)) return false;
      cur.pop_back();
    }
    return true;
  };
  return loop(0);
}

// One solve of FitLMVMAlgo (hipgisaxs_fit_lmvm.cpp:43-180, 190-224): bounded limited-memory BFGS
// from P.x0. The gradient comes from one tangent pass (HipGISAXS::fit_gradient: forward-mode
// derivatives through the ensemble kernel) where the model allows; the remaining components, and
// every component with P.fd_only, by the reference's central differences over +-step/2 (step, or
// 0.04 where the step is not positive), divided by the same step (the reference divides by 0.04
// whatever the step). Projected search directions with a backtracking Armijo line search; stops
// when the projected gradient or the relative decrease falls below the tolerance, or after 200
// iterations.
bool run_lmvm(Problem& P, const HigFitAlgorithm& a, std::vector<double>& best, double& fbest, std::string&) {
  const int d = (int)P.lo.size(), mem = 6, max_iter = 200;
  const double tol = a.tolerance > 0 ? a.tolerance : 1e-6;
  std::vector<double> x = P.x0;
  auto project = [&](std::vector<double>& y) {
    for (int j = 0; j < d; ++j) y[j] = std::min(std::max(y[j], P.lo[j]), P.hi[j]);
  };
  project(x);
  auto grad = [&](const std::vector<double>& y, double& f, std::vector<double>& g) -> bool {
    g.assign(d, 0.0);
    std::vector<int> done(d, 0);
    if (P.fd_only) {
      if (!P.eval(y, f)) return false;
    } else {
      ++P.evals;
      ++P.grad_passes;
      if (!P.sim.fit_gradient(y.data(), P.metric, P.reg, &f, g.data(), done.data())) return false;
    }
    for (int j = 0; j < d; ++j) {
      if (done[j]) continue;
      const double h = P.step[j] > DBL_MIN ? P.step[j] : 0.04;
      std::vector<double> yp = y, ym = y;
      yp[j] += 0.5 * h;
      ym[j] -= 0.5 * h;
      double fp, fm;
      if (!P.eval(yp, fp) || !P.eval(ym, fm)) return false;
      g[j] = (fp - fm) / h;
    }
    return true;
  };
  double f;
  std::vector<double> g;
  if (!grad(x, f, g)) return false;
  std::vector<std::vector<double>> S, Y;
  for (int it = 0; it < max_iter; ++it) {
    // projected gradient: zero where a bound is active and the gradient pushes outwards
    double pg = 0.0;
    std::vector<bool> act(d, false);
    for (int j = 0; j < d; ++j) {
      act[j] = (x[j] <= P.lo[j] && g[j] > 0) || (x[j] >= P.hi[j] && g[j] < 0);
      if (!act[j]) pg = std::max(pg, std::fabs(g[j]));
    }
    if (pg <= tol) break;
    // two-loop recursion on the free variables
    std::vector<double> q(d);
    for (int j = 0; j < d; ++j) q[j] = act[j] ? 0.0 : g[j];
    const int m = (int)S.size();
    std::vector<double> al(m), rho(m);
    for (int k = m - 1; k >= 0; --k) {
      double sy = 0, sq = 0;
      for (int j = 0; j < d; ++j) {
        sy += S[k][j] * Y[k][j];
        sq += S[k][j] * q[j];
      }
      rho[k] = 1.0 / sy;
      al[k] = rho[k] * sq;
      for (int j = 0; j < d; ++j) q[j] -= al[k] * Y[k][j];
    }
    double gamma = 1.0;
    if (m > 0) {
      double sy = 0, yy = 0;
      for (int j = 0; j < d; ++j) {
        sy += S[m - 1][j] * Y[m - 1][j];
        yy += Y[m - 1][j] * Y[m - 1][j];
      }
      gamma = sy / yy;
    } else {
      // first step: 1% of the largest range along the steepest descent
      double r = 0, gn = 0;
      for (int j = 0; j < d; ++j) {
        r = std::max(r, P.hi[j] - P.lo[j]);
        gn = std::max(gn, std::fabs(q[j]));
      }
      gamma = gn > 0 ? 0.01 * r / gn : 1.0;
    }
    for (int j = 0; j < d; ++j) q[j] *= gamma;
    for (int k = 0; k < m; ++k) {
      double yq = 0;
      for (int j = 0; j < d; ++j) yq += Y[k][j] * q[j];
      const double b = rho[k] * yq;
      for (int j = 0; j < d; ++j) q[j] += S[k][j] * (al[k] - b);
    }
    double gd = 0;
    for (int j = 0; j < d; ++j) {
      if (act[j]) q[j] = 0.0;
      gd -= g[j] * q[j];
    }
    if (!(gd < 0)) {  // not a descent direction: restart from steepest descent
      S.clear();
      Y.clear();
      for (int j = 0; j < d; ++j) q[j] = act[j] ? 0.0 : g[j];
      gd = 0;
      for (int j = 0; j < d; ++j) gd -= g[j] * q[j];
      double r = 0, gn = 0;
      for (int j = 0; j < d; ++j) {
        r = std::max(r, P.hi[j] - P.lo[j]);
        gn = std::max(gn, std::fabs(q[j]));
      }
      const double sc = gn > 0 ? 0.01 * r / gn : 1.0;
      for (int j = 0; j < d; ++j) q[j] *= sc;
      gd *= sc;
    }
    double t = 1.0, fn = f;
    std::vector<double> xn(d);
    bool ok = false;
    for (int ls = 0; ls < 30; ++ls) {
      for (int j = 0; j < d; ++j) xn[j] = x[j] - t * q[j];
      project(xn);
      double dec = 0;
      for (int j = 0; j < d; ++j) dec += g[j] * (xn[j] - x[j]);
      if (!P.eval(xn, fn)) return false;
      if (fn <= f + 1e-4 * dec) {
        ok = true;
        break;
      }
      t *= 0.5;
    }
    if (!ok) break;
    std::vector<double> gn;
    double fchk;
    if (!grad(xn, fchk, gn)) return false;
    std::vector<double> s(d), y(d);
    double sy = 0;
    for (int j = 0; j < d; ++j) {
      s[j] = xn[j] - x[j];
      y[j] = gn[j] - g[j];
      sy += s[j] * y[j];
    }
    if (sy > 1e-300) {
      S.push_back(s);
      Y.push_back(y);
      if ((int)S.size() > mem) {
        S.erase(S.begin());
        Y.erase(Y.begin());
      }
    }
    const double rel = std::fabs(f - fchk) / std::max(std::fabs(f), 1e-300);
    x = xn;
    f = fchk;
    g = gn;
    if (rel <= tol) break;
  }
  best = x;
  fbest = f;
  return true;
}

}  // namespace

bool fit_run(HipGISAXS& sim, int algo_index, uint64_t seed, std::vector<double>& best, double& fbest,
             int64_t& evals, std::string& err) {
  const HigInput& in = sim.input();
  if (algo_index < 0 || algo_index >= (int)in.algorithms.size()) {
    err = "no fitting algorithm " + std::to_string(algo_index);
    return false;
  }
  const HigFitAlgorithm& a = in.algorithms[algo_index];
  const int metric = metric_of(a.metric);
  if (metric < 0) {
    err = "unsupported distance metric '" + a.metric + "'";
    return false;
  }
  if (!sim.fit_ready() && !sim.fit_init()) {
    err = sim.error();
    return false;
  }
  Problem P{sim, metric, a.regularization, {}, {}, {}, {}};
  P.fd_only = (sim.flags() & HGX_HIG_FIT_FD_GRADIENT) != 0;
  for (const auto& p : in.fit_params) {
    P.lo.push_back(p.second.min);
    P.hi.push_back(p.second.max);
    P.step.push_back(p.second.step);
    P.x0.push_back(p.second.init);
  }
  if (P.lo.empty()) {
    err = "the input has no fit parameters";
    return false;
  }
  sim.fit_history().clear();
  bool ok;
  if (a.name == "pso") {
    P.reg = 0.0;  // the PSO driver sets no regularization
    ok = run_pso(P, a, seed, best, fbest, err);
  } else if (a.name == "bruteforce") {
    P.reg = 0.0;
    ok = run_bruteforce(P, best, fbest, err);
  } else if (a.name == "lmvm") {
    // FitLMVMAlgo::run's regularization continuation (hipgisaxs_fit_lmvm.cpp:77-172): up to 10 solves,
    // the factor divided by 5 after each, each solve starting from the previous solution; a zero
    // factor gives one solve
    double reg = a.regularization;
    ok = true;
    for (int r = 0; r < 10 && ok; ++r) {
      P.reg = reg;
      ok = run_lmvm(P, a, best, fbest, err);
      if (ok) P.x0 = best;
      reg /= 5.0;
      if (reg <= DBL_MIN) break;
    }
  } else if (a.name == "none_pounders") {  // the objective at the initial values, no fit
    best = P.x0;
    P.reg = 0.0;
    ok = P.eval(best, fbest);
  } else if (a.name == "pounders") {
    err = "POUNDERS (PETSc/TAO) is not implemented; use lmvm, pso or bruteforce";
    ok = false;
  } else {
    err = "unknown optimization algorithm '" + a.name + "'";
    ok = false;
  }
  evals = P.evals;
  if (!ok && err.empty()) err = sim.error();
  return ok;
}

}  // namespace hig

// ----- C ABI ---------------------------------------------------------------------------------------

namespace {

int fit_fail(hgx_hig* h, const std::string& msg, int code = HGX_ERR_INVALID_ARG) {
  h->err = msg;
  return code;
}

void copy_str(const std::string& s, char* out, size_t cap) {
  if (!out || cap == 0) return;
  const size_t n = std::min(s.size(), cap - 1);
  std::memcpy(out, s.data(), n);
  out[n] = 0;
}

}  // namespace

extern "C" {

int hgx_hig_update_param(hgx_hig* h, const char* variable, double value) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!variable) return fit_fail(h, "update_param: null variable");
    if (!h->sim.update_param(variable, value)) return fit_fail(h, h->sim.error());
    return HGX_OK;
  });
}

int hgx_hig_fit_info(hgx_hig* h, int* nparams, int* nalgorithms) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!nparams || !nalgorithms) return fit_fail(h, "fit_info: null output");
    *nparams = (int)h->sim.input().fit_params.size();
    *nalgorithms = (int)h->sim.input().algorithms.size();
    return HGX_OK;
  });
}

int hgx_hig_fit_param(hgx_hig* h, int i, char* key, size_t key_cap, double* range4) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    const auto& fp = h->sim.input().fit_params;
    if (i < 0 || i >= (int)fp.size() || !range4) return fit_fail(h, "fit_param: bad index or output");
    auto it = fp.begin();
    std::advance(it, i);
    copy_str(it->second.key, key, key_cap);
    range4[0] = it->second.min;
    range4[1] = it->second.max;
    range4[2] = it->second.step;
    range4[3] = it->second.init;
    return HGX_OK;
  });
}

int hgx_hig_fit_algorithm(hgx_hig* h, int i, char* name, size_t name_cap, int* metric, double* tol_reg) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    const auto& al = h->sim.input().algorithms;
    if (i < 0 || i >= (int)al.size() || !metric || !tol_reg) return fit_fail(h, "fit_algorithm: bad index or output");
    copy_str(al[i].name, name, name_cap);
    *metric = hig::metric_of(al[i].metric);
    tol_reg[0] = al[i].tolerance;
    tol_reg[1] = al[i].regularization;
    return HGX_OK;
  });
}

int hgx_hig_fit_init(hgx_hig* h) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!h->sim.fit_init()) return fit_fail(h, h->sim.error(), HGX_ERR_INVALID_ARG);
    return HGX_OK;
  });
}

int hgx_hig_fit_objective(hgx_hig* h, int n, const double* x, int metric, double reg, double* value) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!x || !value || n != (int)h->sim.input().fit_params.size())
      return fit_fail(h, "fit_objective: x must hold one value per fit parameter");
    if (metric < HGX_METRIC_CHI2 || metric > HGX_METRIC_CBRT_C_L2) return fit_fail(h, "fit_objective: unknown metric");
    if (!h->sim.fit_ready() && !h->sim.fit_init()) return fit_fail(h, h->sim.error(), HGX_ERR_INVALID_ARG);
    if (!h->sim.fit_objective(x, metric, reg, value)) return fit_fail(h, h->sim.error(), HGX_ERR_INVALID_ARG);
    return HGX_OK;
  });
}

int hgx_hig_fit_gradient(hgx_hig* h, int n, const double* x, int metric, double reg, double* value, double* grad,
                         int* computed) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!x || !value || !grad || !computed) return fit_fail(h, "fit_gradient: null argument");
    if (n != (int)h->sim.input().fit_params.size())
      return fit_fail(h, "fit_gradient: x must hold one value per fit parameter");
    if (metric < HGX_METRIC_CHI2 || metric > HGX_METRIC_CBRT_C_L2) return fit_fail(h, "fit_gradient: unknown metric");
    if (!h->sim.fit_ready() && !h->sim.fit_init()) return fit_fail(h, h->sim.error(), HGX_ERR_INVALID_ARG);
    if (!h->sim.fit_gradient(x, metric, reg, value, grad, computed))
      return fit_fail(h, h->sim.error(), HGX_ERR_INVALID_ARG);
    return HGX_OK;
  });
}

int hgx_hig_fit_run(hgx_hig* h, int algorithm, unsigned long long seed, double* x_best, double* value,
                    long long* evals) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!x_best || !value || !evals) return fit_fail(h, "fit_run: null output");
    std::vector<double> best;
    double f = 0.0;
    int64_t ne = 0;
    std::string err;
    const bool ok = hig::fit_run(h->sim, algorithm, (uint64_t)seed, best, f, ne, err);
    *evals = (long long)ne;
    if (!ok) return fit_fail(h, err, HGX_ERR_INVALID_ARG);
    if (best.size() != h->sim.input().fit_params.size())
      return fit_fail(h, "fit_run: the algorithm evaluated no point", HGX_ERR_INVALID_ARG);
    std::copy(best.begin(), best.end(), x_best);
    *value = f;
    return HGX_OK;
  });
}

int hgx_hig_fit_history(hgx_hig* h, long long cap_rows, double* rows, long long* nrows) {
  return hgx_hig_guard(h, [&]() -> int {
    if (!h) return HGX_ERR_BAD_HANDLE;
    if (!nrows) return fit_fail(h, "fit_history: null output");
    auto& hist = h->sim.fit_history();
    *nrows = (long long)hist.size();
    if (!rows) return HGX_OK;
    const size_t w = h->sim.input().fit_params.size() + 1;
    const long long n = std::min<long long>(cap_rows, (long long)hist.size());
    for (long long r = 0; r < n; ++r) {
      if (hist[r].size() != w) return fit_fail(h, "fit_history: rows of another parameter count", HGX_ERR_INVALID_ARG);
      std::copy(hist[r].begin(), hist[r].end(), rows + r * w);
    }
    return HGX_OK;
  });
}

}  // extern "C"
