// hig_input.hpp -- reader of HipGISAXS input (.hig) files: a generic tokenizer/parser into a tree,
// then the typed model the simulation loop needs (shapes, unit cells, layers, structures, beam,
// output region). Mirrors HiGInput (include/config/hig_input.hpp, src/config/hig_input.cpp:62-2832)
// in its keywords, defaults and unit conversions; fitting blocks are accepted and ignored.
#pragma once

#include <array>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace hig {

// One value of the file: number, string, array or object. Objects keep their fields in file order
// and allow repeated keys (several `shape = {...}` or `param = {...}` blocks).
struct HigNode {
  enum Kind { kNumber, kString, kArray, kObject };
  Kind kind = kNumber;
  double num = 0.0;
  std::string str;
  std::vector<HigNode> items;
  std::vector<std::pair<std::string, HigNode>> fields;
  int line = 0;
};

// Parses text: a sequence of `key = value` assignments (separated by optional commas), where a
// value is a number, a quoted string, a bare word, `[ values ]` or `{ assignments }`; `#` starts a
// comment to the end of the line. Returns false and sets err (with the line) on a syntax error.
bool parse_hig_text(const std::string& text, HigNode& root, std::string& err);

typedef std::array<double, 3> vec3;

// Shape parameter (include/model/shape.hpp:30-95): stat codes as hgx_param_distribution takes them
// (0 none/single, 1 uniform, 2 gaussian); mean = p1, deviation = p2.
struct HigShapeParam {
  std::string type;
  int stat = 0;
  double min = 0.0, max = 0.0, p1 = 0.0, p2 = 0.0;
  int nvalues = 1;
};

struct HigShape {
  std::string key, name;
  int shape = -1;        // HGX_SHAPE_* for analytic shapes, -1 for a mesh file
  std::string file;      // mesh file for custom shapes (name ending .obj / .dat / .hd5 / .shp)
  double xrot = 0.0, yrot = 0.0, zrot = 0.0;  // radians (the file gives degrees)
  double delta = 0.0, beta = 0.0;
  vec3 origin{0.0, 0.0, 0.0};
  std::map<std::string, HigShapeParam> params;  // by type name, the last block of a type wins
};

struct HigUnitcell {
  std::string key;
  std::map<std::string, std::vector<vec3>> elements;  // shape key -> locations
};

struct HigLayer {
  std::string key;
  int order = 0;
  double delta = 0.0, beta = 0.0, thickness = 0.0;
};

struct HigRotation {
  char axis = 'x';  // 'x', 'y', 'z' or 'n' (none)
  double amin = 0.0, amax = 0.0;
  double mean = 0.0, sd = 0.0, location = 0.0, scale = 0.0;
};

struct HigStructure {
  std::string key;
  double iratio = 1.0;
  std::string unitcell_key;
  std::string layer_key = "air";
  std::string grain_type;  // "" (lattice), "paracrystal" or "percusyevick"
  // lattice (src/model/structure.cpp:63-72 defaults)
  std::string lattice_type = "cubic";
  std::string hkl = "100";
  bool abc_set = false;
  vec3 a{1.0, 0.0, 0.0}, b{0.0, 1.0, 0.0}, c{0.0, 0.0, 1.0};
  vec3 transvec{0.0, 0.0, 0.0};
  vec3 repetition{1.0, 1.0, 1.0};
  vec3 scaling{1.0, 1.0, 1.0};  // grain scaling means
  // Grain scaling distribution per lattice axis (GrainScaling, src/model/structure.cpp:35-46): the
  // statistic (none, gaussian or cauchy), its width and sample count. Set by `a|b|c = { stat, mean,
  // stddev, nsamples }` inside the grain's lattice or scaling block.
  enum { kStatNone = 0, kStatGaussian = 1, kStatCauchy = 2, kStatOther = 3 };
  int scale_stat[3] = {kStatNone, kStatNone, kStatNone};
  vec3 scale_sd{0.0, 0.0, 0.0};
  int scale_nvals[3] = {40, 40, 40};
  bool scaling_dist = false;  // some axis has a statistic other than none
  // Grain repetition distribution (GrainRepetitions, include/model/structure.hpp:281-330): per axis
  // unset (-1), none (0, the minimum), range (1) or another statistic (2), with its integer bounds.
  bool repetition_dist = false;
  int rep_stat[3] = {-1, -1, -1};
  double rep_min[3] = {0.0, 0.0, 0.0}, rep_max[3] = {0.0, 0.0, 0.0};
  // ensemble (src/model/structure.cpp:385-390, 236-246 defaults)
  vec3 spacing{0.0, 0.0, 0.0};
  vec3 maxgrains{1.0, 1.0, 1.0};
  std::string distribution = "regular";
  std::string orient_stat = "single";
  HigRotation rot[3];
  HigStructure() {
    rot[0].axis = 'x';
    rot[1].axis = 'y';
    rot[2].axis = 'z';
  }
};

// src/model/inst_scattering.cpp:27-38 defaults; angles in degrees as in the file.
struct HigScattering {
  std::string expt = "gisaxs";
  double alphai_min = 0.1, alphai_max = 0.1, alphai_step = 0.1;
  double phi_min = 0.0, phi_max = 0.0, phi_step = 0.0;
  double tilt_min = 0.0, tilt_max = 0.0, tilt_step = 0.0;
  double photon = 10000.0;
  std::string unit = "ev";
  double smearing = 1.0;
};

// src/model/compute_params.cpp:26-39 defaults.
struct HigCompute {
  std::string pathprefix = ".";
  std::string runname = "run";
  std::string method = "dwba";
  std::string region = "qspace";
  double minpoint[2] = {-1.0, -1.0};
  double maxpoint[2] = {-1.0, -1.0};
  int resolution[2] = {1, 1};  // [columns, rows]
  int nslices = 0;
  std::string structcorr = "nGnE";
  bool saveff = false, savesf = false;
};

struct HigDetector {
  double pixelsize = 0.172, sdd = 4000.0;
  double directbeam[2] = {0.0, 0.0};
  int totalpixels[2] = {1475, 1679};
  std::string origin = "bottom-left";
};

// The fitting block (src/config/hig_input.cpp:302-340, 849-890, 1444-1477, 1524-1571, 1755-1777;
// include/model/fitting_params.hpp, include/config/temp_helpers.hpp).
struct HigFitParam {
  std::string key, variable;  // variable: a path such as shape['s']:param['radius']:p1
  double min = 0.0, max = 0.0, step = -1.0, init = 0.0;
};

struct HigFitReference {
  std::string path, mask;
  std::string region = "qspace";
  double regmin[2] = {0.0, 0.0}, regmax[2] = {0.0, 0.0};
  int npoints[2] = {0, 0};  // parallel, perpendicular
};

struct HigFitAlgorithm {
  std::string name;             // lmvm, pounders, pso, bruteforce, none_pounders
  int order = 0;
  double tolerance = 0.0, regularization = 0.0;
  std::string metric;           // distancemetric
  bool restart = false;
  std::map<std::string, double> params;  // algoparam type -> pvalue
};

class HigInput {
 public:
  // Reads a .hig file; mesh file names of custom shapes are resolved against its directory.
  bool construct_input(const std::string& path);
  // Reads .hig text; base_dir resolves relative mesh file names.
  bool construct_input_text(const std::string& text, const std::string& base_dir = ".");
  const std::string& error() const { return err_; }

  std::map<std::string, HigShape> shapes;
  std::map<std::string, HigUnitcell> unitcells;
  std::map<int, HigLayer> layers;  // by order; -1 is the substrate
  std::map<std::string, HigStructure> structures;
  HigScattering scattering;
  HigCompute compute;
  HigDetector detector;
  // fitting: parameters by key (the order of the fit vector), reference data, algorithms in file order
  std::map<std::string, HigFitParam> fit_params;
  HigFitReference reference;
  bool reference_set = false;
  std::vector<HigFitAlgorithm> algorithms;

  // HiGInput::update_params for one parameter (src/config/hig_input.cpp:2712-2832): sets the model
  // value that `variable` names. False, with the reason in error(), for a malformed or unknown path.
  bool update_param(const std::string& variable, double value);

 private:
  bool fail(const HigNode& at, const std::string& msg);
  bool read_root(const HigNode& root);
  bool read_shape(const HigNode& n);
  bool read_unitcell(const HigNode& n);
  bool read_layer(const HigNode& n);
  bool read_structure(const HigNode& n);
  bool read_grain(const HigNode& n, HigStructure& s);
  bool scaling_axis(const HigNode& n, int ax, HigStructure& s);
  bool read_ensemble(const HigNode& n, HigStructure& s);
  bool read_instrumentation(const HigNode& n);
  bool read_computation(const HigNode& n);
  bool read_fitting(const HigNode& n);
  bool number(const HigNode& n, double& out);
  bool vector(const HigNode& n, size_t count, double* out);
  bool string(const HigNode& n, std::string& out);
  bool refindex(const HigNode& n, double& delta, double& beta);
  std::string err_;
  std::string base_dir_ = ".";
};

// Mesh files of custom shapes (ObjectShapeReader, src/file/objectshape_reader.cpp:52-300): Wavefront
// OBJ vertices and faces ("f i", "f i/j", "f i//k", "f i/j/k", negative indices counting back from
// the last vertex) into 9 doubles per triangle (v1, v2, v3). Polygons of n > 3 vertices are split
// into the fan (v0, v_k, v_k+1); the reference keeps triangles only and drops the rest (*polygons
// counts the faces split).
// Lattice::construct_vectors: the unit-scaled vectors of the structure's lattice (its a, b, c when
// the file gives them).
bool lattice_vectors(const HigStructure& s, vec3& a, vec3& b, vec3& c, std::string& err);

bool read_obj_mesh(const std::string& path, std::vector<double>& tri9, std::string& err, int* polygons = nullptr);
// Raw shape files (RawShapeReader, src/file/rawshape_reader.cpp:39-62): 7 numbers per triangle
// (area, nx, ny, nz, cx, cy, cz) into 8-double approximate-method records (pad 0).
bool read_raw_mesh(const std::string& path, std::vector<double>& rec8, std::string& err);

}  // namespace hig
