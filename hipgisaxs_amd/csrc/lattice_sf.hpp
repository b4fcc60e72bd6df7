// lattice_sf.hpp -- FP64 finite-lattice structure factor of the rotated q-vector (the reference's
// src/sf/sf.cpp:165-245 in closed Dirichlet form), shared by the SF kernel and the fused kernels.
#pragma once

#include "dual.hpp"

namespace hgx {

constexpr double kPiD = 3.141592653589793238462643383279502884;

// ----- structure factor -------------------------------------------------------------------------
//
// The reference (src/sf/sf.cpp:197-245) evaluates, per lattice direction with x = l.q,
//     pow(e, (1-N)/2) * (1 - pow(e, N)) / (1 - e),   e = exp(i x),
// with the principal-branch pow of std::complex, and N where |1 - e| vanishes. pow(e, p) is
// exp(i p x') with x' = wrap(Re x) + i Im x (wrap into (-pi, pi]), so the product is the
// Dirichlet kernel D_N(x') = sin(N x'/2) / sin(x'/2), evaluated here in that closed form, which
// keeps full relative accuracy next to the Bragg peaks where the ratio form cancels.
// For an integer N up to 48 (a lattice's repetition count, the same on every lane) the same value is
// U_{N-1}(cos h), h = x'/2, the Chebyshev polynomial of the second kind, by its three-term recurrence
// (stable: for |cos h| > 1 it follows the growing solution): one complex cosine and N - 2 complex
// multiply-adds in place of two complex sines and a division, the peak N exact.
__device__ inline cd dirichlet(cd x, double N) {
  double w = x.r - 2.0 * kPiD * rint(x.r * (0.5 / kPiD));
  cd xp = cd{w, x.i};
  cd h = 0.5 * xp;
  if (N <= 48.0 && N == rint(N)) {
    const int n = (int)N;
    if (n <= 1) return cd{1.0, 0.0};
    const cd c2 = 2.0 * ccos_d(h);
    cd a = cd{1.0, 0.0}, b = c2;  // U_0, U_1
    for (int k = 2; k < n; ++k) {
      const cd t = c2 * b - a;
      a = b;
      b = t;
    }
    return b;
  }
  cd den = csin_d(h);
  if (cabs_d(den) < 1e-300) return cd{N, 0.0};
  return cdiv(csin_d(N * h), den);
}

template <class R>
struct SFArgsT {
  R la[3], lb[3], lc[3];
  double N[3];
  R center[3];
  R dw;  // sigma^2 / 2
  int compat_lb2;
};
using SFArgs = SFArgsT<double>;

// D_N on duals: the plain value, the derivative d/dh (ddirichlet_dh) times dh = dx/2 (0 at a peak)
__device__ inline dz dirichlet(dz x, double N) {
  const cd v = dirichlet(x.v, N);
  const double w = x.v.r - 2.0 * kPiD * rint(x.v.r * (0.5 / kPiD));
  const cd h = 0.5 * cd{w, x.v.i};
  if (cabs_d(csin_d(h)) < 1e-300) return dz{v, cd{0.0, 0.0}};
  return dz{v, ddirichlet_dh(h, N) * (0.5 * x.d)};
}

__device__ inline bool nonzero(double x) { return x != 0.0; }
__device__ inline bool nonzero(dd x) { return x.v != 0.0 || x.d != 0.0; }

// S(q) for one rotated q-vector: product of the three Dirichlet kernels, the grain-centre phase
// and the Debye-Waller damping.
template <class C, class R>
__device__ inline C sf_value(const C mq[3], const SFArgsT<R>& a) {
  C xa = mq[0] * a.la[0] + mq[1] * a.la[1] + mq[2] * a.la[2];
  C xb = mq[0] * a.lb[0] + mq[1] * a.lb[1] + (a.compat_lb2 ? mkc(a.lb[2] * a.lb[2], R{}) : mq[2] * a.lb[2]);
  C xc = mq[0] * a.lc[0] + mq[1] * a.lc[1] + mq[2] * a.lc[2];
  C s = dirichlet(xa, a.N[0]) * dirichlet(xb, a.N[1]) * dirichlet(xc, a.N[2]);
  s = s * cexpi_d(mq[0] * a.center[0] + mq[1] * a.center[1] + mq[2] * a.center[2]);
  if (nonzero(a.dw)) s = exp(-a.dw * (norm2(mq[0]) + norm2(mq[1]) + norm2(mq[2]))) * s;
  return s;
}

}  // namespace hgx
