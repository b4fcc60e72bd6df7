// adjoint_core.hpp -- assembly of the approximate-triangle record adjoint from the product-grid
// bilinear forms (shared by adjoint.hip and grating.hip).
#pragma once

#include "hgx_internal.hpp"

namespace hgx {

// part: [6 forms][nyt][ntri] complex sums of f(z) g(y) Wt Ez Ey over one y-tile each, forms
// B11, B1y, B1yy, Bz1, Bzy, Bzz. With V = (qx0 B11, B1y, Bz1) and
// M = [[qx0^2 B11, qx0 B1y, qx0 Bz1], [qx0 B1y, B1yy, Bzy], [qx0 Bz1, Bzy, Bzz]]:
// dL/dA = Re V, dL/dc = Re i M A (frame of the records in rec).
__device__ inline void adj_assemble(const double* __restrict__ part, int nyt, int64_t ntri, int64_t t,
                                    const double* __restrict__ rec, double qx0, double gA[3], double gc[3]) {
  cd B[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    B[f] = cd{0.0, 0.0};
    for (int y = 0; y < nyt; ++y) {
      const double* o = part + 2 * (((int64_t)f * nyt + y) * ntri + t);
      B[f] = B[f] + cd{o[0], o[1]};
    }
  }
  const cd V[3] = {qx0 * B[0], B[1], B[3]};
  const cd M[3][3] = {{(qx0 * qx0) * B[0], qx0 * B[1], qx0 * B[3]},
                      {qx0 * B[1], B[2], B[4]},
                      {qx0 * B[3], B[4], B[5]}};
  const double* r = rec + 8 * t;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    gA[i] = V[i].r;
    cd ma = M[i][0] * r[0] + M[i][1] * r[1] + M[i][2] * r[2];
    gc[i] = -ma.i;  // Re(i z) = -Im z
  }
}

}  // namespace hgx
