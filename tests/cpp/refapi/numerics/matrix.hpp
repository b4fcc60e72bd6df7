// Test stand-in for the reference's include/numerics/matrix.hpp: RotMatrix_t with private storage
// and the public constructors and rotate() of the reference (the adapters reach the entries through
// rotate() only).
#pragma once

#include <cmath>
#include <vector>

#include <common/globals.hpp>
#include <common/typedefs.hpp>

namespace hig {
class RotMatrix_t {
 private:
  real_t data_[9];

 public:
  RotMatrix_t() {
    for (int i = 0; i < 9; ++i) data_[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
  RotMatrix_t(int axis, real_t a) {
    const real_t c = std::cos(a), s = std::sin(a);
    const real_t x[9] = {1, 0, 0, 0, c, -s, 0, s, c};
    const real_t y[9] = {c, 0, -s, 0, 1, 0, s, 0, c};
    const real_t z[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
    const real_t* m = axis == 0 ? x : axis == 1 ? y : z;
    for (int i = 0; i < 9; ++i) data_[i] = m[i];
  }
  // the test's way in: a matrix from nine row-major entries (vector3_t rows, as the reference's
  // third constructor)
  RotMatrix_t(vector3_t& a, vector3_t& b, vector3_t& c) {
    for (int k = 0; k < 3; ++k) {
      data_[k] = a[k];
      data_[3 + k] = b[k];
      data_[6 + k] = c[k];
    }
  }
  std::vector<complex_t> rotate(real_t x, real_t y, complex_t z) {
    std::vector<complex_t> res(3);
    res[0] = data_[0] * x + data_[1] * y + data_[2] * z;
    res[1] = data_[3] * x + data_[4] * y + data_[5] * z;
    res[2] = data_[6] * x + data_[7] * y + data_[8] * z;
    return res;
  }
};
}  // namespace hig
