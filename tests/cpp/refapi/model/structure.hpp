// Test stand-in for the reference's include/model/structure.hpp: a Lattice with the accessors
// StructureFactorG::compute reads (a(), b(), c(), t()) and setters for the test.
#pragma once

#include <common/globals.hpp>

namespace hig {
class Lattice {
 private:
  vector3_t a_, b_, c_, t_;

 public:
  vector3_t a() { return a_; }
  vector3_t b() { return b_; }
  vector3_t c() { return c_; }
  vector3_t t() { return t_; }
  void a(vector3_t v) { a_ = v; }
  void b(vector3_t v) { b_ = v; }
  void c(vector3_t v) { c_ = v; }
};
}  // namespace hig
