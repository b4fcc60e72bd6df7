// hull.cpp -- 3-D convex hull (quickhull with conflict lists) for the mesh generators: the convex
// polyhedra of the batched-shapes workload are hulls of points on ellipsoids.
//
// Start from the tetrahedron of four extreme points; every other point sits in the outside set of
// the first face it lies above. Repeatedly take a face with a non-empty outside set, its farthest
// point (the eye), the connected set of faces the eye sees, and the horizon (edges between seen and
// unseen faces); replace the seen faces by the fan from the horizon to the eye and hand their
// outside points to the new faces. Faces are kept counter-clockwise seen from outside, with the face
// across each edge, so the horizon walk needs no search.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "hgx.h"
#include "hgx_hig.h"

namespace {

struct Face {
  int v[3];
  double n[3], d;  // unit outward normal, n.x = d on the plane
  int nb[3];       // face across edge i = (v[i], v[i+1])
  std::vector<int> out;
  bool alive = true;
  int seen = -1;   // eye round that marked it visible
};

struct Hull {
  const double* p;
  std::vector<Face> f;
  double eps;

  double dist(const Face& F, int i) const {
    const double* x = p + 3 * (size_t)i;
    return F.n[0] * x[0] + F.n[1] * x[1] + F.n[2] * x[2] - F.d;
  }
  bool plane(Face& F) const {
    const double *a = p + 3 * (size_t)F.v[0], *b = p + 3 * (size_t)F.v[1], *c = p + 3 * (size_t)F.v[2];
    const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, w[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
    const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (!(l > 0)) return false;
    for (int k = 0; k < 3; ++k) F.n[k] = n[k] / l;
    F.d = F.n[0] * a[0] + F.n[1] * a[1] + F.n[2] * a[2];
    return true;
  }
};

inline uint64_t edge_key(int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; }

}  // namespace

extern "C" int64_t hgx_hig_convex_hull(int64_t n, const double* pts, int64_t cap, int* tri) {
  if (n < 4 || !pts || n > (int64_t)1 << 30) return HGX_ERR_INVALID_ARG;
  Hull H;
  H.p = pts;
  double scale = 0.0;
  for (int64_t i = 0; i < 3 * n; ++i) {
    if (!std::isfinite(pts[i])) return HGX_ERR_INVALID_ARG;
    scale = std::max(scale, std::fabs(pts[i]));
  }
  H.eps = 1e-12 * std::max(scale, 1e-300) * 8;
  auto P = [&](int64_t i) { return pts + 3 * (size_t)i; };
  // the initial tetrahedron: extremes along the axis of largest spread, the farthest point from
  // their line, the farthest from their plane
  int i0 = 0, i1 = 0;
  double best = -1.0;
  for (int ax = 0; ax < 3; ++ax) {
    int lo = 0, hi = 0;
    for (int64_t i = 1; i < n; ++i) {
      if (P(i)[ax] < P(lo)[ax]) lo = (int)i;
      if (P(i)[ax] > P(hi)[ax]) hi = (int)i;
    }
    if (P(hi)[ax] - P(lo)[ax] > best) {
      best = P(hi)[ax] - P(lo)[ax];
      i0 = lo;
      i1 = hi;
    }
  }
  if (!(best > H.eps)) return HGX_ERR_INVALID_ARG;
  int i2 = -1;
  best = 0.0;
  {
    const double* a = P(i0);
    const double u[3] = {P(i1)[0] - a[0], P(i1)[1] - a[1], P(i1)[2] - a[2]};
    for (int64_t i = 0; i < n; ++i) {
      const double w[3] = {P(i)[0] - a[0], P(i)[1] - a[1], P(i)[2] - a[2]};
      const double c[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
      const double d = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
      if (d > best) {
        best = d;
        i2 = (int)i;
      }
    }
  }
  if (i2 < 0) return HGX_ERR_INVALID_ARG;
  Face base;
  base.v[0] = i0;
  base.v[1] = i1;
  base.v[2] = i2;
  if (!H.plane(base)) return HGX_ERR_INVALID_ARG;
  int i3 = -1;
  best = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double d = std::fabs(H.dist(base, (int)i));
    if (d > best) {
      best = d;
      i3 = (int)i;
    }
  }
  if (i3 < 0 || !(best > H.eps)) return HGX_ERR_INVALID_ARG;  // all points coplanar
  const int tv[4] = {i0, i1, i2, i3};
  double centre[3] = {0, 0, 0};
  for (int k = 0; k < 4; ++k)
    for (int j = 0; j < 3; ++j) centre[j] += 0.25 * P(tv[k])[j];
  std::unordered_map<uint64_t, int> owner;  // directed edge -> face holding it
  owner.reserve((size_t)n * 8);
  auto add_face = [&](int a, int b, int c) -> int {
    Face F;
    F.v[0] = a;
    F.v[1] = b;
    F.v[2] = c;
    H.plane(F);
    H.f.push_back(F);
    return (int)H.f.size() - 1;
  };
  {
    const int faces[4][3] = {{i0, i1, i2}, {i0, i3, i1}, {i1, i3, i2}, {i2, i3, i0}};
    for (auto& t : faces) {
      int a = t[0], b = t[1], c = t[2];
      Face F;
      F.v[0] = a;
      F.v[1] = b;
      F.v[2] = c;
      H.plane(F);
      if (F.n[0] * centre[0] + F.n[1] * centre[1] + F.n[2] * centre[2] - F.d > 0) std::swap(b, c);
      add_face(a, b, c);
    }
    for (int fi = 0; fi < 4; ++fi)
      for (int e = 0; e < 3; ++e) owner[edge_key(H.f[fi].v[e], H.f[fi].v[(e + 1) % 3])] = fi;
    for (int fi = 0; fi < 4; ++fi)
      for (int e = 0; e < 3; ++e) H.f[fi].nb[e] = owner.at(edge_key(H.f[fi].v[(e + 1) % 3], H.f[fi].v[e]));
  }
  for (int64_t i = 0; i < n; ++i) {
    if (i == i0 || i == i1 || i == i2 || i == i3) continue;
    for (int fi = 0; fi < 4; ++fi)
      if (H.dist(H.f[fi], (int)i) > H.eps) {
        H.f[fi].out.push_back((int)i);
        break;
      }
  }
  std::vector<int> stack = {0, 1, 2, 3}, visible, todo, fresh;
  std::unordered_map<int, int> new_by_start, new_by_end;
  int round = 0;
  while (!stack.empty()) {
    const int f0 = stack.back();
    stack.pop_back();
    if (!H.f[f0].alive || H.f[f0].out.empty()) continue;
    // the eye: the farthest outside point of f0
    int eye = -1;
    double far = -1.0;
    for (int i : H.f[f0].out) {
      const double d = H.dist(H.f[f0], i);
      if (d > far) {
        far = d;
        eye = i;
      }
    }
    // the faces the eye sees (connected, from f0)
    ++round;
    visible.clear();
    todo.assign(1, f0);
    H.f[f0].seen = round;
    while (!todo.empty()) {
      const int fi = todo.back();
      todo.pop_back();
      visible.push_back(fi);
      for (int e = 0; e < 3; ++e) {
        const int g = H.f[fi].nb[e];
        if (H.f[g].seen == round) continue;
        if (H.dist(H.f[g], eye) > H.eps) {
          H.f[g].seen = round;
          todo.push_back(g);
        }
      }
    }
    // horizon edges (a, b) of visible faces whose neighbour is hidden: the new face (a, b, eye)
    fresh.clear();
    new_by_start.clear();
    new_by_end.clear();
    for (int fi : visible)
      for (int e = 0; e < 3; ++e) {
        const int g = H.f[fi].nb[e];
        if (H.f[g].seen == round) continue;
        const int a = H.f[fi].v[e], b = H.f[fi].v[(e + 1) % 3];
        const int nf = add_face(a, b, eye);
        H.f[nf].nb[0] = g;
        for (int k = 0; k < 3; ++k)
          if (H.f[g].v[k] == b && H.f[g].v[(k + 1) % 3] == a) H.f[g].nb[k] = nf;
        new_by_start[a] = nf;
        new_by_end[b] = nf;
        fresh.push_back(nf);
      }
    // the fan's own edges: (b, eye) meets the new face starting at b, (eye, a) the one ending at a
    for (int nf : fresh) {
      Face& F = H.f[nf];
      auto s = new_by_start.find(F.v[1]);
      auto t = new_by_end.find(F.v[0]);
      if (s == new_by_start.end() || t == new_by_end.end()) return HGX_ERR_INVALID_ARG;  // broken horizon
      F.nb[1] = s->second;
      F.nb[2] = t->second;
    }
    // outside points of the seen faces go to the first new face they lie above
    for (int fi : visible) {
      Face& V = H.f[fi];
      V.alive = false;
      for (int i : V.out) {
        if (i == eye) continue;
        for (int nf : fresh)
          if (H.dist(H.f[nf], i) > H.eps) {
            H.f[nf].out.push_back(i);
            break;
          }
      }
      std::vector<int>().swap(V.out);
    }
    for (int nf : fresh)
      if (!H.f[nf].out.empty()) stack.push_back(nf);
  }
  int64_t nt = 0;
  for (const Face& F : H.f)
    if (F.alive) {
      if (tri && nt < cap)
        for (int k = 0; k < 3; ++k) tri[3 * nt + k] = F.v[k];
      ++nt;
    }
  return nt;
}
