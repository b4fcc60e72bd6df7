// Test stand-in for the reference's include/common/globals.hpp: vector3_t as the reference defines
// its storage and element access (vec_[3], operator[]).
#pragma once

#include <common/typedefs.hpp>

namespace hig {
typedef struct vector3_t {
  real_t vec_[3];
  vector3_t() { vec_[0] = vec_[1] = vec_[2] = 0; }
  vector3_t(real_t a, real_t b, real_t c) {
    vec_[0] = a;
    vec_[1] = b;
    vec_[2] = c;
  }
  real_t& operator[](int i) { return vec_[i]; }
} vector3_t;
}  // namespace hig
