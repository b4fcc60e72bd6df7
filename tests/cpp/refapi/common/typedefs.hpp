// Test stand-in for the reference's include/common/typedefs.hpp (DOUBLEP, GPU build): only the
// names hipgisaxs_amd/host/hig_gpu.hpp and tests/cpp/hig_adapter_shim.cpp use, with the reference's
// definitions (real_t, complex_t, the vectors, cucomplex_t as two doubles x, y, triangle_t).
#pragma once

#include <complex>
#include <vector>

namespace hig {
typedef double real_t;
typedef struct {
  double x, y;
} cucomplex_t;  // cuDoubleComplex in the reference's CUDA build
typedef std::complex<real_t> complex_t;
typedef std::vector<real_t> real_vec_t;
typedef std::vector<complex_t> complex_vec_t;
typedef struct Triangle {
  real_t v1[3];
  real_t v2[3];
  real_t v3[3];
} triangle_t;
}  // namespace hig
