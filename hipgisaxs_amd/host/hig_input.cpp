// hig_input.cpp -- .hig tokenizer/parser and the typed model (see hig_input.hpp).
#include "hig_input.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "hgx.h"

namespace hig {

namespace {

const double kPi = 3.14159265358979323846;

// ----- tokenizer ---------------------------------------------------------------------------------

struct Token {
  enum Type { kEnd, kPunct, kString, kNumber, kWord };
  Type type = kEnd;
  char punct = 0;
  std::string text;
  double num = 0.0;
  int line = 1;
};

class Lexer {
 public:
  explicit Lexer(const std::string& s) : s_(s) {}
  bool next(Token& t, std::string& err) {
    skip();
    t = Token();
    t.line = line_;
    if (p_ >= s_.size()) return true;
    char ch = s_[p_];
    if (ch == '{' || ch == '}' || ch == '[' || ch == ']' || ch == '=' || ch == ',') {
      t.type = Token::kPunct;
      t.punct = ch;
      ++p_;
      return true;
    }
    if (ch == '"' || ch == '\'') {
      size_t e = s_.find(ch, p_ + 1);
      if (e == std::string::npos) {
        err = "unterminated string at line " + std::to_string(line_);
        return false;
      }
      t.type = Token::kString;
      t.text = s_.substr(p_ + 1, e - p_ - 1);
      for (char c : t.text)
        if (c == '\n') ++line_;
      p_ = e + 1;
      return true;
    }
    bool numeric = std::isdigit((unsigned char)ch) || ch == '.' ||
                   ((ch == '-' || ch == '+') && p_ + 1 < s_.size() &&
                    (std::isdigit((unsigned char)s_[p_ + 1]) || s_[p_ + 1] == '.'));
    if (numeric) {
      const char* b = s_.c_str() + p_;
      char* e = nullptr;
      double v = std::strtod(b, &e);
      if (e == b) {
        err = "bad number at line " + std::to_string(line_);
        return false;
      }
      t.type = Token::kNumber;
      t.num = v;
      p_ += (size_t)(e - b);
      return true;
    }
    if (std::isalpha((unsigned char)ch) || ch == '_') {
      size_t b = p_;
      while (p_ < s_.size()) {
        char c = s_[p_];
        if (std::isalnum((unsigned char)c) || c == '_' || c == ':' || c == '.' || c == '-') ++p_;
        else break;
      }
      t.type = Token::kWord;
      t.text = s_.substr(b, p_ - b);
      return true;
    }
    err = std::string("unexpected character '") + ch + "' at line " + std::to_string(line_);
    return false;
  }

 private:
  void skip() {
    while (p_ < s_.size()) {
      char c = s_[p_];
      if (c == '#') {
        while (p_ < s_.size() && s_[p_] != '\n') ++p_;
      } else if (c == '\n') {
        ++line_;
        ++p_;
      } else if (std::isspace((unsigned char)c)) {
        ++p_;
      } else {
        break;
      }
    }
  }
  const std::string& s_;
  size_t p_ = 0;
  int line_ = 1;
};

class Parser {
 public:
  explicit Parser(const std::string& s) : lex_(s) {}
  bool run(HigNode& root, std::string& err) {
    if (!advance(err)) return false;
    root = HigNode();
    root.kind = HigNode::kObject;
    root.line = 1;
    return assignments(root, 0, err, 0);
  }

 private:
  bool advance(std::string& err) { return lex_.next(tok_, err); }
  bool is_punct(char c) const { return tok_.type == Token::kPunct && tok_.punct == c; }
  bool where(std::string& err, const std::string& msg) {
    err = msg + " at line " + std::to_string(tok_.line);
    return false;
  }
  // key = value pairs until `close` ('}' or end of input when close == 0)
  bool assignments(HigNode& obj, char close, std::string& err, int depth) {
    if (depth > 64) return where(err, "nesting too deep");
    while (true) {
      if (close == 0 ? tok_.type == Token::kEnd : is_punct(close)) return true;
      if (tok_.type == Token::kEnd) return where(err, std::string("missing '") + close + "'");
      if (tok_.type != Token::kWord) return where(err, "expected a keyword");
      std::string key = tok_.text;
      int line = tok_.line;
      if (!advance(err)) return false;
      if (!is_punct('=')) return where(err, "expected '=' after '" + key + "'");
      if (!advance(err)) return false;
      HigNode v;
      if (!value(v, err, depth + 1)) return false;
      v.line = line;
      obj.fields.emplace_back(key, std::move(v));
      if (is_punct(',')) {
        if (!advance(err)) return false;
      }
    }
  }
  bool value(HigNode& v, std::string& err, int depth) {
    if (depth > 64) return where(err, "nesting too deep");
    v.line = tok_.line;
    if (is_punct('{')) {
      v.kind = HigNode::kObject;
      if (!advance(err)) return false;
      if (!assignments(v, '}', err, depth)) return false;
      return advance(err);
    }
    if (is_punct('[')) {
      v.kind = HigNode::kArray;
      if (!advance(err)) return false;
      while (!is_punct(']')) {
        if (tok_.type == Token::kEnd) return where(err, "missing ']'");
        HigNode item;
        if (!value(item, err, depth + 1)) return false;
        v.items.push_back(std::move(item));
        if (is_punct(',')) {
          if (!advance(err)) return false;
        }
      }
      return advance(err);
    }
    if (tok_.type == Token::kNumber) {
      v.kind = HigNode::kNumber;
      v.num = tok_.num;
      return advance(err);
    }
    if (tok_.type == Token::kString || tok_.type == Token::kWord) {
      v.kind = HigNode::kString;
      v.str = tok_.text;
      return advance(err);
    }
    return where(err, "expected a value");
  }
  Lexer lex_;
  Token tok_;
};

bool has_extension(const std::string& s, const char* ext) {
  size_t n = std::char_traits<char>::length(ext);
  return s.size() > n && s.compare(s.size() - n, n, ext) == 0;
}

// StatisticType names (include/config/token_mapper.hpp:314-321) for shape parameters
bool shape_stat(const std::string& s, int& code) {
  if (s == "none" || s == "single") code = 0;
  else if (s == "uniform") code = 1;
  else if (s == "gaussian") code = 2;
  else return false;
  return true;
}

}  // namespace

bool parse_hig_text(const std::string& text, HigNode& root, std::string& err) {
  Parser p(text);
  return p.run(root, err);
}

// ----- typed model -------------------------------------------------------------------------------

bool HigInput::fail(const HigNode& at, const std::string& msg) {
  err_ = "line " + std::to_string(at.line) + ": " + msg;
  return false;
}

bool HigInput::number(const HigNode& n, double& out) {
  if (n.kind != HigNode::kNumber) return fail(n, "expected a number");
  out = n.num;
  return true;
}

bool HigInput::vector(const HigNode& n, size_t count, double* out) {
  if (n.kind != HigNode::kArray || n.items.size() != count)
    return fail(n, "expected a vector of " + std::to_string(count) + " numbers");
  for (size_t i = 0; i < count; ++i)
    if (!number(n.items[i], out[i])) return false;
  return true;
}

bool HigInput::string(const HigNode& n, std::string& out) {
  if (n.kind != HigNode::kString) return fail(n, "expected a string");
  out = n.str;
  return true;
}

bool HigInput::refindex(const HigNode& n, double& delta, double& beta) {
  if (n.kind != HigNode::kObject) return fail(n, "refindex must be an object");
  for (const auto& f : n.fields) {
    if (f.first == "delta") {
      if (!number(f.second, delta)) return false;
    } else if (f.first == "beta") {
      if (!number(f.second, beta)) return false;
    } else {
      return fail(f.second, "unknown keyword '" + f.first + "' in refindex");
    }
  }
  return true;
}

bool HigInput::construct_input(const std::string& path) {
  std::ifstream f(path);
  if (!f) {
    err_ = "cannot open " + path;
    return false;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  size_t slash = path.find_last_of('/');
  return construct_input_text(ss.str(), slash == std::string::npos ? "." : path.substr(0, slash));
}

bool HigInput::construct_input_text(const std::string& text, const std::string& base_dir) {
  *this = HigInput();
  base_dir_ = base_dir;
  HigNode root;
  if (!parse_hig_text(text, root, err_)) return false;
  return read_root(root);
}

bool HigInput::read_root(const HigNode& root) {
  const HigNode* top = nullptr;
  for (const auto& f : root.fields) {
    if (f.first == "hipGisaxsInput") {
      if (top) return fail(f.second, "more than one hipGisaxsInput block");
      top = &f.second;
    } else {
      return fail(f.second, "unknown top-level keyword '" + f.first + "'");
    }
  }
  if (!top || top->kind != HigNode::kObject) {
    err_ = "no hipGisaxsInput = { ... } block";
    return false;
  }
  for (const auto& f : top->fields) {
    const std::string& k = f.first;
    bool ok = true;
    if (k == "shape") ok = read_shape(f.second);
    else if (k == "unitcell") ok = read_unitcell(f.second);
    else if (k == "layer") ok = read_layer(f.second);
    else if (k == "structure") ok = read_structure(f.second);
    else if (k == "instrumentation") ok = read_instrumentation(f.second);
    else if (k == "computation") ok = read_computation(f.second);
    else if (k == "fitting") ok = read_fitting(f.second);
    else return fail(f.second, "unknown keyword '" + k + "'");
    if (!ok) return false;
  }
  if (layers.count(0) == 0) {  // HiGInput::construct_layer_profile inserts the vacuum layer
    HigLayer vac;
    vac.key = "vacuum";
    layers[0] = vac;
  }
  return true;
}

bool HigInput::read_shape(const HigNode& n) {
  if (n.kind != HigNode::kObject) return fail(n, "shape must be an object");
  HigShape s;
  for (const auto& f : n.fields) {
    const std::string& k = f.first;
    const HigNode& v = f.second;
    if (k == "key") {
      if (!string(v, s.key)) return false;
    } else if (k == "name") {
      if (!string(v, s.name)) return false;
      static const std::map<std::string, int> names = {
          {"sphere", HGX_SHAPE_SPHERE}, {"cylinder", HGX_SHAPE_CYLINDER}, {"box", HGX_SHAPE_BOX},
          {"cube", HGX_SHAPE_CUBE},     {"prism3", HGX_SHAPE_PRISM3},     {"prism6", HGX_SHAPE_PRISM6},
          {"pyramid", HGX_SHAPE_PYRAMID}};
      auto it = names.find(s.name);
      if (it != names.end()) {
        s.shape = it->second;
      } else if (has_extension(s.name, ".obj") || has_extension(s.name, ".dat") ||
                 has_extension(s.name, ".hd5") || has_extension(s.name, ".shp")) {
        s.shape = -1;  // shape_custom (include/config/token_mapper.hpp:60-69)
        s.file = (!s.name.empty() && s.name[0] == '/') ? s.name : base_dir_ + "/" + s.name;
      } else if (s.name == "hcylinder" || s.name == "randcylinders" || s.name == "trunccone" ||
                 s.name == "prism3x" || s.name == "sawtooth") {
        s.shape = -2;  // a reference shape name this engine has no form factor for
      } else {
        return fail(v, "unknown shape name '" + s.name + "'");
      }
    } else if (k == "originvec") {
      if (!vector(v, 3, s.origin.data())) return false;
    } else if (k == "xrot" || k == "yrot" || k == "zrot") {
      double deg;
      if (!number(v, deg)) return false;
      double r = deg * kPi / 180.;  // src/config/hig_input.cpp:1100-1109
      (k == "xrot" ? s.xrot : k == "yrot" ? s.yrot : s.zrot) = r;
    } else if (k == "refindex") {
      if (!refindex(v, s.delta, s.beta)) return false;
    } else if (k == "param") {
      if (v.kind != HigNode::kObject) return fail(v, "param must be an object");
      HigShapeParam p;
      for (const auto& g : v.fields) {
        if (g.first == "type") {
          if (!string(g.second, p.type)) return false;
          if (p.type != "radius" && p.type != "xsize" && p.type != "ysize" && p.type != "height" &&
              p.type != "edge" && p.type != "baseangle")
            return fail(g.second, "unknown shape parameter type '" + p.type + "'");
        } else if (g.first == "stat") {
          std::string st;
          if (!string(g.second, st)) return false;
          if (!shape_stat(st, p.stat))
            return fail(g.second, "shape parameter statistic '" + st + "' is not supported (none, uniform, gaussian)");
        } else if (g.first == "min") {
          if (!number(g.second, p.min)) return false;
        } else if (g.first == "max") {
          if (!number(g.second, p.max)) return false;
        } else if (g.first == "p1" || g.first == "mean") {
          if (!number(g.second, p.p1)) return false;
        } else if (g.first == "p2" || g.first == "stddev") {
          if (!number(g.second, p.p2)) return false;
        } else if (g.first == "nvalues") {
          double nv;
          if (!number(g.second, nv)) return false;
          if (nv < 1 || nv > 4096) return fail(g.second, "nvalues must be in [1, 4096]");
          p.nvalues = (int)nv;
        } else {
          return fail(g.second, "unknown keyword '" + g.first + "' in param");
        }
      }
      if (p.type.empty()) return fail(v, "param without a type");
      s.params[p.type] = p;
    } else {
      return fail(v, "unknown keyword '" + k + "' in shape");
    }
  }
  if (s.key.empty()) return fail(n, "shape without a key");
  if (s.name.empty()) return fail(n, "shape '" + s.key + "' without a name");
  shapes[s.key] = s;
  return true;
}

bool HigInput::read_unitcell(const HigNode& n) {
  if (n.kind != HigNode::kObject) return fail(n, "unitcell must be an object");
  HigUnitcell u;
  for (const auto& f : n.fields) {
    if (f.first == "key") {
      if (!string(f.second, u.key)) return false;
    } else if (f.first == "element") {
      const HigNode& e = f.second;
      if (e.kind != HigNode::kObject) return fail(e, "element must be an object");
      std::string skey;
      std::vector<vec3> locs;
      for (const auto& g : e.fields) {
        if (g.first == "shape:key") {
          if (!string(g.second, skey)) return false;
        } else if (g.first == "locations") {
          const HigNode& l = g.second;
          if (l.kind != HigNode::kArray || l.items.empty()) return fail(l, "locations must be a list of vectors");
          for (const auto& it : l.items) {
            vec3 p;
            if (!vector(it, 3, p.data())) return false;
            locs.push_back(p);
          }
        } else {
          return fail(g.second, "unknown keyword '" + g.first + "' in element");
        }
      }
      if (skey.empty()) return fail(e, "element without shape:key");
      if (locs.empty()) return fail(e, "element without locations");
      u.elements[skey] = locs;  // a repeated shape key replaces the earlier list, as the reference's map does
    } else {
      return fail(f.second, "unknown keyword '" + f.first + "' in unitcell");
    }
  }
  if (u.key.empty()) return fail(n, "unitcell without a key");
  unitcells[u.key] = u;
  return true;
}

bool HigInput::read_layer(const HigNode& n) {
  if (n.kind != HigNode::kObject) return fail(n, "layer must be an object");
  HigLayer l;
  for (const auto& f : n.fields) {
    if (f.first == "key") {
      if (!string(f.second, l.key)) return false;
    } else if (f.first == "order") {
      double o;
      if (!number(f.second, o)) return false;
      l.order = (int)o;
    } else if (f.first == "thickness") {
      if (!number(f.second, l.thickness)) return false;
    } else if (f.first == "refindex") {
      if (!refindex(f.second, l.delta, l.beta)) return false;
    } else {
      return fail(f.second, "unknown keyword '" + f.first + "' in layer");
    }
  }
  if (l.order < -1) return fail(n, "layer order must be >= -1");
  layers[l.order] = l;
  return true;
}

// One axis of a grain scaling distribution: { stat, mean, stddev, nsamples } (the mean / stddev /
// nsamples / stat tokens under a lattice axis, src/config/hig_input.cpp:1206-1263, 1642-1652).
bool HigInput::scaling_axis(const HigNode& n, int ax, HigStructure& s) {
  for (const auto& f : n.fields) {
    if (f.first == "stat") {
      std::string st;
      if (!string(f.second, st)) return false;
      s.scale_stat[ax] = (st == "none" || st == "single") ? HigStructure::kStatNone
                         : st == "gaussian"              ? HigStructure::kStatGaussian
                         : st == "cauchy"                ? HigStructure::kStatCauchy
                                                         : HigStructure::kStatOther;
    } else if (f.first == "mean") {
      if (!number(f.second, s.scaling[ax])) return false;
    } else if (f.first == "stddev") {
      if (!number(f.second, s.scale_sd[ax])) return false;
    } else if (f.first == "nsamples") {
      double x;
      if (!number(f.second, x)) return false;
      if (x < 1 || x > 100000 || x != std::floor(x)) return fail(f.second, "nsamples must be a positive integer");
      s.scale_nvals[ax] = (int)x;
    } else {
      return fail(f.second, "unknown keyword '" + f.first + "' in a scaling distribution");
    }
  }
  s.scaling_dist = false;
  for (int i = 0; i < 3; ++i) s.scaling_dist = s.scaling_dist || s.scale_stat[i] != HigStructure::kStatNone;
  return true;
}

bool HigInput::read_grain(const HigNode& n, HigStructure& s) {
  if (n.kind != HigNode::kObject) return fail(n, "grain must be an object");
  for (const auto& f : n.fields) {
    const std::string& k = f.first;
    const HigNode& v = f.second;
    if (k == "unitcell:key") {
      if (!string(v, s.unitcell_key)) return false;
    } else if (k == "layer:key") {
      if (!string(v, s.layer_key)) return false;
    } else if (k == "type") {
      if (!string(v, s.grain_type)) return false;
    } else if (k == "lattice") {
      if (v.kind != HigNode::kObject) return fail(v, "lattice must be an object");
      for (const auto& g : v.fields) {
        if (g.first == "type") {
          if (!string(g.second, s.lattice_type)) return false;
          if (s.lattice_type != "bcc" && s.lattice_type != "cubic" && s.lattice_type != "fcc" &&
              s.lattice_type != "fco" && s.lattice_type != "hcp" && s.lattice_type != "hex")
            return fail(g.second, "unknown lattice type '" + s.lattice_type + "'");
        } else if (g.first == "hkl") {
          if (g.second.kind == HigNode::kNumber) {
            s.hkl = std::to_string((long long)g.second.num);
          } else if (!string(g.second, s.hkl)) {
            return false;
          }
        } else if ((g.first == "a" || g.first == "b" || g.first == "c") && g.second.kind == HigNode::kObject) {
          if (!scaling_axis(g.second, g.first[0] - 'a', s)) return false;
        } else if (g.first == "a" || g.first == "b" || g.first == "c") {
          vec3& t = g.first == "a" ? s.a : g.first == "b" ? s.b : s.c;
          if (!vector(g.second, 3, t.data())) return false;
          s.abc_set = true;
        } else if (g.first == "abangle" || g.first == "caratio") {
          double unused;  // Lattice::construct_vectors reads neither (src/model/structure.cpp:86-182)
          if (!number(g.second, unused)) return false;
        } else {
          return fail(g.second, "unknown keyword '" + g.first + "' in lattice");
        }
      }
    } else if (k == "repetition") {
      if (!vector(v, 3, s.repetition.data())) return false;
    } else if (k == "repetitiondist") {
      // { xrepetition = { stat, min, max }, y..., z... } (src/config/hig_input.cpp:248-262, 924-990)
      if (v.kind != HigNode::kObject) return fail(v, "repetitiondist must be an object");
      s.repetition_dist = true;
      for (const auto& r : v.fields) {
        const int ax = r.first == "xrepetition" ? 0 : r.first == "yrepetition" ? 1 : r.first == "zrepetition" ? 2 : -1;
        if (ax < 0) return fail(r.second, "unknown keyword '" + r.first + "' in repetitiondist");
        if (r.second.kind != HigNode::kObject) return fail(r.second, r.first + " must be an object");
        for (const auto& q : r.second.fields) {
          if (q.first == "stat") {
            std::string st;
            if (!string(q.second, st)) return false;
            s.rep_stat[ax] = (st == "none" || st == "single") ? 0 : st == "range" ? 1 : 2;
          } else if (q.first == "min" || q.first == "max") {
            double x;
            if (!number(q.second, x)) return false;
            if (x < 0 || x > 4294967295.0) return fail(q.second, "repetition bounds must be non-negative integers");
            (q.first == "min" ? s.rep_min : s.rep_max)[ax] = std::floor(x);  // unsigned in the reference
          } else if (q.first != "mean" && q.first != "stddev") {
            return fail(q.second, "unknown keyword '" + q.first + "' in " + r.first);
          }
        }
      }
    } else if (k == "scaling") {
      if (v.kind == HigNode::kNumber) {
        s.scaling = {v.num, v.num, v.num};
      } else if (v.kind == HigNode::kArray) {
        if (!vector(v, 3, s.scaling.data())) return false;
      } else if (v.kind == HigNode::kObject) {
        for (const auto& a : v.fields) {
          if ((a.first != "a" && a.first != "b" && a.first != "c") || a.second.kind != HigNode::kObject)
            return fail(a.second, "scaling distributions are given per axis: a|b|c = { stat, mean, stddev, nsamples }");
          if (!scaling_axis(a.second, a.first[0] - 'a', s)) return false;
        }
      } else {
        return fail(v, "scaling can be a scalar, vector[3] or a distribution");
      }
    } else if (k == "transvec") {
      if (!vector(v, 3, s.transvec.data())) return false;
    } else if (k == "dimensions" || k == "xspacing" || k == "yspacing" || k == "domain" || k == "volfraction") {
      // paracrystal / Percus-Yevick parameters: recorded by grain_type
    } else {
      return fail(v, "unknown keyword '" + k + "' in grain");
    }
  }
  return true;
}

bool HigInput::read_ensemble(const HigNode& n, HigStructure& s) {
  if (n.kind != HigNode::kObject) return fail(n, "ensemble must be an object");
  for (const auto& f : n.fields) {
    const std::string& k = f.first;
    const HigNode& v = f.second;
    if (k == "spacing") {
      if (!vector(v, 3, s.spacing.data())) return false;
    } else if (k == "maxgrains") {
      if (!vector(v, 3, s.maxgrains.data())) return false;
      for (double m : s.maxgrains)
        if (m < 1 || m > 1e6 || m != std::floor(m)) return fail(v, "maxgrains must be positive integers");
    } else if (k == "distribution") {
      if (!string(v, s.distribution)) return false;
    } else if (k == "orientations") {
      if (v.kind != HigNode::kObject) return fail(v, "orientations must be an object");
      for (const auto& g : v.fields) {
        if (g.first == "stat") {
          if (!string(g.second, s.orient_stat)) return false;
        } else if (g.first == "rot1" || g.first == "rot2" || g.first == "rot3") {
          HigRotation& r = s.rot[g.first[3] - '1'];
          if (g.second.kind != HigNode::kObject) return fail(g.second, g.first + " must be an object");
          for (const auto& h : g.second.fields) {
            if (h.first == "axis") {
              std::string ax;
              if (!string(h.second, ax)) return false;
              if (ax != "x" && ax != "y" && ax != "z" && ax != "n")
                return fail(h.second, "rotation axis must be x, y, z or n");
              r.axis = ax[0];
            } else if (h.first == "angles") {
              const HigNode& a = h.second;
              if (a.kind != HigNode::kArray || a.items.empty() || a.items.size() > 2)
                return fail(a, "angles must be [min] or [min max]");
              if (!number(a.items[0], r.amin)) return false;
              r.amax = r.amin;
              if (a.items.size() == 2 && !number(a.items[1], r.amax)) return false;
            } else if (h.first == "anglemean") {
              if (!number(h.second, r.mean)) return false;
            } else if (h.first == "anglesd") {
              if (!number(h.second, r.sd)) return false;
            } else if (h.first == "anglelocation") {
              if (!number(h.second, r.location)) return false;
            } else if (h.first == "anglescale") {
              if (!number(h.second, r.scale)) return false;
            } else {
              return fail(h.second, "unknown keyword '" + h.first + "' in " + g.first);
            }
          }
        } else {
          return fail(g.second, "unknown keyword '" + g.first + "' in orientations");
        }
      }
    } else {
      return fail(v, "unknown keyword '" + k + "' in ensemble");
    }
  }
  return true;
}

bool HigInput::read_structure(const HigNode& n) {
  if (n.kind != HigNode::kObject) return fail(n, "structure must be an object");
  HigStructure s;
  for (const auto& f : n.fields) {
    if (f.first == "key") {
      if (!string(f.second, s.key)) return false;
    } else if (f.first == "iratio") {
      if (!number(f.second, s.iratio)) return false;
      if (!(s.iratio > 0)) return fail(f.second, "iratio must be positive");  // hig_input.cpp:1158-1163
    } else if (f.first == "grain") {
      if (!read_grain(f.second, s)) return false;
    } else if (f.first == "ensemble") {
      if (!read_ensemble(f.second, s)) return false;
    } else if (f.first == "paracrystal" || f.first == "percusyevick") {
      s.grain_type = f.first;
    } else {
      return fail(f.second, "unknown keyword '" + f.first + "' in structure");
    }
  }
  if (s.key.empty()) return fail(n, "structure without a key");
  structures[s.key] = s;
  return true;
}

bool HigInput::read_instrumentation(const HigNode& n) {
  if (n.kind != HigNode::kObject) return fail(n, "instrumentation must be an object");
  for (const auto& f : n.fields) {
    const HigNode& v = f.second;
    if (f.first == "scattering") {
      if (v.kind != HigNode::kObject) return fail(v, "scattering must be an object");
      for (const auto& g : v.fields) {
        const std::string& k = g.first;
        const HigNode& w = g.second;
        if (k == "expt") {
          if (!string(w, scattering.expt)) return false;
          if (scattering.expt != "gisaxs" && scattering.expt != "saxs")
            return fail(w, "expt must be gisaxs or saxs");
        } else if (k == "alphai" || k == "inplanerot" || k == "tilt") {
          double* r = k == "alphai" ? &scattering.alphai_min : k == "inplanerot" ? &scattering.phi_min
                                                                                 : &scattering.tilt_min;
          if (w.kind != HigNode::kObject) return fail(w, k + " must be an object");
          for (const auto& h : w.fields) {
            int idx = h.first == "min" ? 0 : h.first == "max" ? 1 : h.first == "step" ? 2 : -1;
            if (idx < 0) return fail(h.second, "unknown keyword '" + h.first + "' in " + k);
            if (!number(h.second, r[idx])) return false;
            // a step given for alphai also sets the in-plane rotation step: the reference's case
            // falls through (src/config/hig_input.cpp:1028-1032)
            if (k == "alphai" && idx == 2) scattering.phi_step = r[idx];
          }
        } else if (k == "photon") {
          if (w.kind != HigNode::kObject) return fail(w, "photon must be an object");
          for (const auto& h : w.fields) {
            if (h.first == "value") {
              if (!number(h.second, scattering.photon)) return false;
            } else if (h.first == "unit") {
              if (!string(h.second, scattering.unit)) return false;
            } else {
              return fail(h.second, "unknown keyword '" + h.first + "' in photon");
            }
          }
        } else if (k == "smearing") {
          if (!number(w, scattering.smearing)) return false;
        } else if (k == "polarization" || k == "coherence" || k == "spotarea") {
          // recorded by the reference, read by no computation
        } else {
          return fail(w, "unknown keyword '" + k + "' in scattering");
        }
      }
    } else if (f.first == "detector") {
      if (v.kind != HigNode::kObject) return fail(v, "detector must be an object");
      for (const auto& g : v.fields) {
        const std::string& k = g.first;
        if (k == "pixelsize") {
          if (!number(g.second, detector.pixelsize)) return false;
        } else if (k == "sdd") {
          if (!number(g.second, detector.sdd)) return false;
        } else if (k == "directbeam") {
          if (!vector(g.second, 2, detector.directbeam)) return false;
        } else if (k == "totalpixels") {
          double tp[2];
          if (!vector(g.second, 2, tp)) return false;
          detector.totalpixels[0] = (int)tp[0];
          detector.totalpixels[1] = (int)tp[1];
        } else if (k == "origin") {
          if (!string(g.second, detector.origin)) return false;
        } else {
          return fail(g.second, "unknown keyword '" + k + "' in detector");
        }
      }
    } else {
      return fail(v, "unknown keyword '" + f.first + "' in instrumentation");
    }
  }
  return true;
}

bool HigInput::read_computation(const HigNode& n) {
  if (n.kind != HigNode::kObject) return fail(n, "computation must be an object");
  for (const auto& f : n.fields) {
    const std::string& k = f.first;
    const HigNode& v = f.second;
    if (k == "pathprefix") {
      if (!string(v, compute.pathprefix)) return false;
    } else if (k == "runname") {
      if (!string(v, compute.runname)) return false;
    } else if (k == "method") {
      if (!string(v, compute.method)) return false;
    } else if (k == "outputregion") {
      if (v.kind != HigNode::kObject) return fail(v, "outputregion must be an object");
      for (const auto& g : v.fields) {
        if (g.first == "type") {
          if (!string(g.second, compute.region)) return false;
        } else if (g.first == "minpoint") {
          if (!vector(g.second, 2, compute.minpoint)) return false;
        } else if (g.first == "maxpoint") {
          if (!vector(g.second, 2, compute.maxpoint)) return false;
        } else {
          return fail(g.second, "unknown keyword '" + g.first + "' in outputregion");
        }
      }
    } else if (k == "resolution") {
      double r[2];
      if (!vector(v, 2, r)) return false;
      if (r[0] < 2 || r[1] < 2 || r[0] > 65536 || r[1] > 65536) return fail(v, "resolution must be in [2, 65536]");
      compute.resolution[0] = (int)r[0];
      compute.resolution[1] = (int)r[1];
    } else if (k == "nslices") {
      double ns;
      if (!number(v, ns)) return false;
      compute.nslices = (int)ns;
    } else if (k == "structcorrelation") {
      if (!string(v, compute.structcorr)) return false;
    } else if (k == "saveff" || k == "savesf") {
      std::string b;
      if (!string(v, b)) return false;
      bool t = b == "true" || b == "True" || b == "TRUE" || b == "yes" || b == "Yes" || b == "YES";
      (k == "saveff" ? compute.saveff : compute.savesf) = t;
    } else if (k == "palette") {
      std::string unused;
      if (!string(v, unused)) return false;
    } else {
      return fail(v, "unknown keyword '" + k + "' in computation");
    }
  }
  return true;
}

// fitting = { fitparam = {...} ... referencedata = {...} algorithm = {...} ... }: parameters are kept
// by key (the reference's maps order the fit vector by key); a duplicate key or a parameter without a
// key or variable is an error (src/config/hig_input.cpp:307-321).
bool HigInput::read_fitting(const HigNode& n) {
  if (n.kind != HigNode::kObject) return fail(n, "fitting must be an object");
  for (const auto& f : n.fields) {
    const std::string& k = f.first;
    const HigNode& v = f.second;
    if (v.kind != HigNode::kObject) return fail(v, k + " must be an object");
    if (k == "fitparam") {
      HigFitParam p;
      for (const auto& g : v.fields) {
        if (g.first == "key") {
          if (!string(g.second, p.key)) return false;
        } else if (g.first == "variable") {
          if (!string(g.second, p.variable)) return false;
        } else if (g.first == "init") {
          if (!number(g.second, p.init)) return false;
        } else if (g.first == "range") {
          if (g.second.kind != HigNode::kObject) return fail(g.second, "range must be an object");
          for (const auto& r : g.second.fields) {
            double* dst = r.first == "min" ? &p.min : r.first == "max" ? &p.max : r.first == "step" ? &p.step : nullptr;
            if (!dst) return fail(r.second, "unknown keyword '" + r.first + "' in range");
            if (!number(r.second, *dst)) return false;
          }
        } else {
          return fail(g.second, "unknown keyword '" + g.first + "' in fitparam");
        }
      }
      if (p.key.empty() || p.variable.empty()) return fail(v, "incomplete fit parameter definition");
      if (fit_params.count(p.key)) return fail(v, "duplicate key '" + p.key + "' in fit parameters");
      fit_params[p.key] = p;
    } else if (k == "referencedata") {
      HigFitReference& r = reference;
      for (const auto& g : v.fields) {
        if (g.first == "path" || g.first == "mask") {
          std::string s;
          if (!string(g.second, s)) return false;
          if (!s.empty() && s[0] != '/') s = base_dir_ + "/" + s;
          (g.first == "path" ? r.path : r.mask) = s;
        } else if (g.first == "fitregion") {
          if (g.second.kind != HigNode::kObject) return fail(g.second, "fitregion must be an object");
          for (const auto& q : g.second.fields) {
            if (q.first == "type") {
              if (!string(q.second, r.region)) return false;
            } else if (q.first == "regmin") {
              if (!vector(q.second, 2, r.regmin)) return false;
            } else if (q.first == "regmax") {
              if (!vector(q.second, 2, r.regmax)) return false;
            } else {
              return fail(q.second, "unknown keyword '" + q.first + "' in fitregion");
            }
          }
        } else if (g.first == "npoints") {
          if (g.second.kind != HigNode::kObject) return fail(g.second, "npoints must be an object");
          for (const auto& q : g.second.fields) {
            double x;
            if (q.first != "parallel" && q.first != "perpendicular")
              return fail(q.second, "unknown keyword '" + q.first + "' in npoints");
            if (!number(q.second, x)) return false;
            r.npoints[q.first == "parallel" ? 0 : 1] = (int)x;
          }
        } else {
          return fail(g.second, "unknown keyword '" + g.first + "' in referencedata");
        }
      }
      reference_set = true;
    } else if (k == "algorithm") {
      HigFitAlgorithm a;
      for (const auto& g : v.fields) {
        double x;
        if (g.first == "algoname") {
          if (!string(g.second, a.name)) return false;
        } else if (g.first == "algoorder") {
          if (!number(g.second, x)) return false;
          a.order = (int)x;
        } else if (g.first == "tolerance") {
          if (!number(g.second, a.tolerance)) return false;
        } else if (g.first == "regularization") {
          if (!number(g.second, a.regularization)) return false;
        } else if (g.first == "distancemetric") {
          if (!string(g.second, a.metric)) return false;
        } else if (g.first == "restart") {
          std::string b;
          if (!string(g.second, b)) return false;
          a.restart = b == "true" || b == "True" || b == "TRUE" || b == "yes" || b == "Yes" || b == "YES";
        } else if (g.first == "algoparam") {
          if (g.second.kind != HigNode::kObject) return fail(g.second, "algoparam must be an object");
          std::string type;
          double value = 0.0;
          for (const auto& q : g.second.fields) {
            if (q.first == "type") {
              if (!string(q.second, type)) return false;
            } else if (q.first == "pvalue") {
              if (!number(q.second, value)) return false;
            } else {
              return fail(q.second, "unknown keyword '" + q.first + "' in algoparam");
            }
          }
          if (type.empty()) return fail(g.second, "algoparam without a type");
          a.params[type] = value;
        } else {
          return fail(g.second, "unknown keyword '" + g.first + "' in algorithm");
        }
      }
      algorithms.push_back(a);
    } else {
      return fail(v, "unknown keyword '" + k + "' in fitting");
    }
  }
  return true;
}

}  // namespace hig

namespace hig {

bool read_obj_mesh(const std::string& path, std::vector<double>& tri9, std::string& err, int* polygons) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open mesh file " + path;
    return false;
  }
  std::vector<std::array<double, 3>> v;
  tri9.clear();
  int npoly = 0, lineno = 0;
  std::string line;
  while (std::getline(f, line)) {
    ++lineno;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ls(line);
    std::string tag;
    if (!(ls >> tag)) continue;
    if (tag == "v") {
      std::array<double, 3> p;
      if (!(ls >> p[0] >> p[1] >> p[2])) {
        err = path + ":" + std::to_string(lineno) + ": bad vertex";
        return false;
      }
      v.push_back(p);
    } else if (tag == "f") {
      std::vector<long> idx;
      std::string tok;
      while (ls >> tok) {
        char* end = nullptr;
        long i = std::strtol(tok.c_str(), &end, 10);
        if (end == tok.c_str() || (*end != '\0' && *end != '/')) {
          err = path + ":" + std::to_string(lineno) + ": bad face index '" + tok + "'";
          return false;
        }
        long n = (long)v.size();
        long k = i > 0 ? i - 1 : n + i;
        if (i == 0 || k < 0 || k >= n) {
          err = path + ":" + std::to_string(lineno) + ": face index out of range";
          return false;
        }
        idx.push_back(k);
      }
      if (idx.size() < 3) {
        err = path + ":" + std::to_string(lineno) + ": face with fewer than 3 vertices";
        return false;
      }
      if (idx.size() > 3) ++npoly;
      for (size_t t = 1; t + 1 < idx.size(); ++t) {
        const long c[3] = {idx[0], idx[t], idx[t + 1]};
        for (long k : c)
          for (int d = 0; d < 3; ++d) tri9.push_back(v[k][d]);
      }
    }
    // vt, vn, g, o, s, l, usemtl, mtllib and comments carry nothing the form factor uses
  }
  if (tri9.empty()) {
    err = "mesh file " + path + " has no faces";
    return false;
  }
  if (polygons) *polygons = npoly;
  return true;
}

bool read_raw_mesh(const std::string& path, std::vector<double>& rec8, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open mesh file " + path;
    return false;
  }
  rec8.clear();
  double r[7];
  while (true) {
    int k = 0;
    while (k < 7 && (f >> r[k])) ++k;
    if (k == 0) break;
    if (k < 7) {
      err = "raw mesh file " + path + ": the number count is not a multiple of 7";
      return false;
    }
    rec8.insert(rec8.end(), r, r + 7);
    rec8.push_back(0.0);
  }
  if (!f.eof()) {
    err = "raw mesh file " + path + ": not a list of numbers";
    return false;
  }
  if (rec8.empty()) {
    err = "mesh file " + path + " has no triangles";
    return false;
  }
  return true;
}

}  // namespace hig
