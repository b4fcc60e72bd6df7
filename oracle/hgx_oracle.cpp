// hgx_oracle.cpp -- FP64 CPU oracle for the GISAXS form-factor / structure-factor / DWBA path.
//
// TEST INFRASTRUCTURE ONLY. This file is the correctness checker and the reported CPU baseline:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The product
// library (hipgisaxs_amd/csrc) never links, loads or calls it.
//
// It restates, in plain FP64 C++ (std::complex<double>, libm, OpenMP over q-points exactly like
// the reference's CPU loop nests), the algorithms of the reference files cited per function
// (paths relative to the reference repository root). The reference itself cannot be compiled
// here (it needs Boost, FFTW, yaml-cpp, MPI, libtiff), so this restatement is the oracle; it is
// pinned by closed forms, invariants and brute force in tests/test_oracle_*.py.
//
// Conventions shared with the product C ABI:
//   * q-grid: qx[npix], qy[npix] real; qz[nq] complex interleaved (re,im), nq = nbranch*npix,
//     q-point i <-> pixel i % npix (src/model/qgrid.cpp:121-124).
//   * every FF / SF output is complex interleaved double[2*nq]; intensities are double[npix].
//   * rot is a row-major 3x3 matrix; the rotated q is mq = rot * (qx, qy, qz)
//     (include/numerics/matrix.hpp:123-130).
//
// Parity pinning status per function is recorded in the header comment of each function.

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <algorithm>

typedef std::complex<double> cplx;

static const double kPi = 3.141592653589793;
static const double kTiny = std::numeric_limits<double>::min();  // TINY_ (include/common/constants.hpp:25)
static const double kCuTiny = 1.0e-13;                            // CUTINY_ (constants.hpp:26)

static inline cplx qz_at(const double* qz, int64_t i) { return cplx(qz[2 * i], qz[2 * i + 1]); }
static inline void store(double* out, int64_t i, cplx v) { out[2 * i] = v.real(); out[2 * i + 1] = v.imag(); }

// mq = R * (qx, qy, qz) with complex z (include/numerics/matrix.hpp:123-130).
static inline void rotate_q(const double* R, double qx, double qy, cplx qz, cplx mq[3]) {
  mq[0] = R[0] * qx + R[1] * qy + R[2] * qz;
  mq[1] = R[3] * qx + R[4] * qy + R[5] * qz;
  mq[2] = R[6] * qx + R[7] * qy + R[8] * qz;
}

static inline int sgn_ref(double a) { return a < 0 ? -1 : 1; }  // include/utils/utilities.hpp:31-34

extern "C" {

// ---------------------------------------------------------------------------------------------
// a1: QGrid::create, region_qspace branch (src/model/qgrid.cpp:30-99).
// Rows are alpha_f (reversed so row 0 has the highest qz, :75), columns are 2theta_f.
// theta uses cos(alpha_[0]) of the *lowest* row, taken before the reversal (:56-58).
// Outputs qx,qy,qz[nrow*ncol] and alpha_f[nrow] (already reversed, = QGrid::alpha(i)).
// Pinned by: tests/test_oracle_qgrid.py (Ewald-sphere invariant |k_f| = k0, inversion of the
// vertical and horizontal q-points, acos clamp).
int hgxo_qgrid_qspace(int ncol, int nrow, double qy_min, double qz_min, double qy_max, double qz_max,
                      double alpha_i, double k0, double* qx, double* qy, double* qz, double* alpha_f) {
  if (ncol < 2 || nrow < 2) return -1;
  std::vector<double> alpha(nrow), theta(ncol);
  double dq = (qz_max - qz_min) / (nrow - 1);
  for (int i = 0; i < nrow; ++i) {
    double qpt = qz_min + i * dq;
    alpha[i] = std::asin(qpt / k0 - std::sin(alpha_i));
  }
  dq = (qy_max - qy_min) / (ncol - 1);
  double cos_ai = std::cos(alpha_i);
  double cos_af = std::cos(alpha[0]);
  for (int j = 0; j < ncol; ++j) {
    double qpt = qy_min + j * dq;
    double kf2 = std::pow(qpt / k0, 2);
    double tmp = (cos_af * cos_af + cos_ai * cos_ai - kf2) / (2 * cos_af * cos_ai);
    if (tmp > 1.0) tmp = 1.0;
    theta[j] = sgn_ref(qpt) * std::acos(tmp);
  }
  std::reverse(alpha.begin(), alpha.end());
  for (int i = 0; i < nrow; ++i) {
    alpha_f[i] = alpha[i];
    for (int j = 0; j < ncol; ++j) {
      double tth = theta[j], alf = alpha[i];
      int64_t p = (int64_t)i * ncol + j;
      qx[p] = k0 * (std::cos(alf) * std::cos(tth) - std::cos(alpha_i));
      qy[p] = k0 * (std::cos(alf) * std::sin(tth));
      qz[p] = k0 * (std::sin(alf) + std::sin(alpha_i));
    }
  }
  return 0;
}

// Explicit rectangular grid: qx = 0, qy uniform in [qy_min, qy_max] over columns, qz uniform in
// [qz_min, qz_max] over rows with row 0 = highest qz (the row order of qgrid.cpp:75).
// alpha_f[row] = asin(qz/k0 - sin alpha_i) as in qgrid.cpp:51. No reference counterpart: this is
// the workload's own grid (the reference always derives qx from the Ewald sphere).
// Pinned by: direct construction in tests (it is only a linspace).
int hgxo_qgrid_explicit(int ncol, int nrow, double qy_min, double qz_min, double qy_max, double qz_max,
                        double alpha_i, double k0, double* qx, double* qy, double* qz, double* alpha_f) {
  if (ncol < 2 || nrow < 2) return -1;
  for (int i = 0; i < nrow; ++i) {
    double qzv = qz_min + (nrow - 1 - i) * ((qz_max - qz_min) / (nrow - 1));
    alpha_f[i] = std::asin(qzv / k0 - std::sin(alpha_i));
    for (int j = 0; j < ncol; ++j) {
      int64_t p = (int64_t)i * ncol + j;
      qx[p] = 0.0;
      qy[p] = qy_min + j * ((qy_max - qy_min) / (ncol - 1));
      qz[p] = qzv;
    }
  }
  return 0;
}

// Flat-detector grid: pixel (row i, col j) sees 2theta = atan((j - col_center) p / D) and
// alpha_f = atan((nrow - 1 - i + 0.5 - row_horizon) p / D) (row 0 = top = highest alpha_f), then the
// Ewald-sphere q of qgrid.cpp:76-84. The reference's own detector helper hard-codes its geometry
// (qgrid.cpp:201-214); this is the workload's explicit geometry.
// Pinned by: tests (|k_f| = k0 on every pixel; symmetric columns give qy -> -qy).
int hgxo_qgrid_detector(int ncol, int nrow, double pixel, double distance, double col_center,
                        double row_horizon, double alpha_i, double k0,
                        double* qx, double* qy, double* qz, double* alpha_f) {
  if (ncol < 1 || nrow < 1 || distance <= 0) return -1;
  for (int i = 0; i < nrow; ++i) {
    double alf = std::atan(((nrow - 1 - i) + 0.5 - row_horizon) * pixel / distance);
    alpha_f[i] = alf;
    for (int j = 0; j < ncol; ++j) {
      double tth = std::atan((j - col_center) * pixel / distance);
      int64_t p = (int64_t)i * ncol + j;
      qx[p] = k0 * (std::cos(alf) * std::cos(tth) - std::cos(alpha_i));
      qy[p] = k0 * (std::cos(alf) * std::sin(tth));
      qz[p] = k0 * (std::sin(alf) + std::sin(alpha_i));
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------------------------
// a2: QGrid::create_qz_extended (src/model/qgrid.cpp:102-127).
// kzi = -k0 sqrt(sin^2 ai - dnl), kzf = sgn(kzf0) sqrt(kzf0^2 - k0^2 dnl), kzf0 = qz - k0 sin ai;
// branches [kzf - kzi, -kzf - kzi, kzf + kzi, -kzf + kzi].
// Pinned by: tests (dnl = 0 gives branch 0 == qz, closed forms of the other three branches).
void hgxo_qz_extended(int64_t npix, const double* qz, double k0, double alpha_i, double dnl_re,
                      double dnl_im, double* qz_ext) {
  cplx dnl(dnl_re, dnl_im);
  double sin_ai = std::sin(alpha_i);
  double kzi_0 = -1 * k0 * sin_ai;
  cplx kzi = -1.0 * k0 * std::sqrt(sin_ai * sin_ai - dnl);
  for (int64_t i = 0; i < npix; ++i) {
    double kzf_0 = qz[i] + kzi_0;
    cplx kzf = (double)sgn_ref(kzf_0) * std::sqrt(kzf_0 * kzf_0 - k0 * k0 * dnl);
    store(qz_ext, i, kzf - kzi);
    store(qz_ext, i + npix, -kzf - kzi);
    store(qz_ext, i + 2 * npix, kzf + kzi);
    store(qz_ext, i + 3 * npix, -kzf + kzi);
  }
}

// ---------------------------------------------------------------------------------------------
// a3: MultiLayer::parratt_recursion and propagation_coeffs (src/model/multilayer.cpp:29-155).
// Stack = [vacuum, layers..., substrate] with one_minus_n2 = 2 delta + 2 i beta
// (include/model/common.hpp:36); interface depth z_i = -sum_{j=1..i} thickness_j (:50-56).
// delta/beta/thickness describe layers 1..nl (the last one is the substrate).
static void parratt(int NL, const cplx* n2, const double* zval, double alpha, double k0, int order,
                    cplx* Tout, cplx* Rout) {
  double sina = std::sin(alpha);
  std::vector<cplx> T(NL, 0.0), R(NL, 0.0), kz(NL, 0.0);
  T.back() = cplx(1.0, 0.0);
  for (int i = 0; i < NL; ++i) kz[i] = -k0 * std::sqrt(sina * sina - n2[i]);
  const cplx mi(0.0, -1.0);
  for (int i = NL - 2; i > -1; --i) {
    cplx pij = (kz[i] + kz[i + 1]) / (2. * kz[i]);
    cplx mij = (kz[i] - kz[i + 1]) / (2. * kz[i]);
    double z = zval[i];
    cplx exp_p = std::exp(mi * (kz[i + 1] + kz[i]) * z);
    cplx exp_m = std::exp(mi * (kz[i + 1] - kz[i]) * z);
    cplx a00 = pij * exp_m, a01 = mij * std::conj(exp_p);
    cplx a10 = mij * exp_p, a11 = pij * std::conj(exp_m);
    T[i] = a00 * T[i + 1] + a01 * R[i + 1];
    R[i] = a10 * T[i + 1] + a11 * R[i + 1];
  }
  cplx t0 = T[0];
  *Tout = T[order] / t0;
  *Rout = R[order] / t0;
}

// Fills fc[4*npix] = [Ti conj(Tf), Ti conj(Rf), Ri conj(Tf), Ri conj(Rf)] with Tf = Rf = 0 for
// alpha_f <= 0 (multilayer.cpp:110-155). nl counts layers below the vacuum (substrate last).
// Pinned by: tests (single interface = textbook Fresnel r,t; |r| = 1 below the critical angle
// when beta = 0; slab = Airy formula; delta = beta = 0 gives T = 1, R = 0).
int hgxo_fresnel(int nl, const double* delta, const double* beta, const double* thickness, int order,
                 double k0, double alpha_i, int nrow, int ncol, const double* alpha_f, double* fc) {
  if (nl < 1 || order < 0 || order > nl - 1) return -1;  // order -1 (substrate) is rejected (:111-116)
  int NL = nl + 1;
  std::vector<cplx> n2(NL);
  std::vector<double> zval(NL);
  n2[0] = 0.0;
  zval[0] = 0.0;
  double z = 0.0;
  for (int i = 1; i < NL; ++i) {
    n2[i] = cplx(2 * delta[i - 1], 2 * beta[i - 1]);
    z -= thickness[i - 1];
    zval[i] = z;
  }
  cplx Ti, Ri;
  parratt(NL, n2.data(), zval.data(), alpha_i, k0, order, &Ti, &Ri);
  std::vector<cplx> Tf(nrow, 0.0), Rf(nrow, 0.0);
#pragma omp parallel for
  for (int i = 0; i < nrow; ++i) {
    if (alpha_f[i] > 0) parratt(NL, n2.data(), zval.data(), alpha_f[i], k0, order, &Tf[i], &Rf[i]);
  }
  int64_t npix = (int64_t)nrow * ncol;
#pragma omp parallel for
  for (int64_t i = 0; i < npix; ++i) {
    int64_t j = i / ncol;
    store(fc, i, Ti * std::conj(Tf[j]));
    store(fc, i + npix, Ti * std::conj(Rf[j]));
    store(fc, i + 2 * npix, Ri * std::conj(Tf[j]));
    store(fc, i + 3 * npix, Ri * std::conj(Rf[j]));
  }
  return 0;
}

// ---------------------------------------------------------------------------------------------
// a7: NumericFormFactorC::compute_approx_triangle (src/ff/cpu/ff_tri_cpu.cpp:200-246).
// F(q) = sum_t (-i)(q.n_t) s_t exp(+i q.c_t) / |q|^2, |q|^2 = sum_j |mq_j|^2 (:235-240).
// rec: 8 doubles per triangle (s, nx, ny, nz, cx, cy, cz, pad) (src/ff/ff_num.cpp:1027-1030).
// Deviation (fixes the reference's 0/0): at |q|^2 == 0 exactly, F = sum_t s_t (n_t.c_t) / 3,
// the q -> 0 limit (= V for a closed mesh).
// Pinned by: F(0) = V, triangulated-box convergence to the analytic box, brute-force volume
// quadrature, F(-q) = conj F(q), rotation and translation invariants (tests/test_oracle_ff.py).
void hgxo_ff_approx(int64_t ntri, const double* rec, int64_t npix, int64_t nq, const double* qx,
                    const double* qy, const double* qz, const double* rot, double* ff) {
  double vol0 = 0.0;
  for (int64_t t = 0; t < ntri; ++t) {
    const double* r = rec + 8 * t;
    vol0 += r[0] * (r[1] * r[4] + r[2] * r[5] + r[3] * r[6]);
  }
  vol0 /= 3.0;
  const cplx nj(0.0, -1.0), np(0.0, 1.0);
#pragma omp parallel for schedule(static)
  for (int64_t iz = 0; iz < nq; ++iz) {
    int64_t iy = iz % npix;
    cplx mq[3];
    rotate_q(rot, qx[iy], qy[iy], qz_at(qz, iz), mq);
    double q2 = std::norm(mq[0]) + std::norm(mq[1]) + std::norm(mq[2]);
    if (q2 == 0.0) { store(ff, iz, cplx(vol0, 0.0)); continue; }
    cplx acc = 0.0;
    for (int64_t t = 0; t < ntri; ++t) {
      const double* r = rec + 8 * t;
      cplx qn = mq[0] * r[1] + mq[1] * r[2] + mq[2] * r[3];
      cplx qt = mq[0] * r[4] + mq[1] * r[5] + mq[2] * r[6];
      acc += (nj * qn * r[0] * std::exp(np * qt) / q2);
    }
    store(ff, iz, acc);
  }
}

// The same sum evaluated the way the reference's CPU loop is written (src/ff/cpu/ff_tri_cpu.cpp:219-242):
// the q-point is rotated again for every (q, triangle) pair into a freshly allocated std::vector
// (:233), i.e. one heap allocation per pair. Timed beside hgxo_ff_approx as the "reference as written"
// CPU baseline; its values equal hgxo_ff_approx's (same arithmetic, q = 0 handled the same way).
void hgxo_ff_approx_as_written(int64_t ntri, const double* rec, int64_t npix, int64_t nq, const double* qx,
                               const double* qy, const double* qz, const double* rot, double* ff) {
  double vol0 = 0.0;
  for (int64_t t = 0; t < ntri; ++t) {
    const double* r = rec + 8 * t;
    vol0 += r[0] * (r[1] * r[4] + r[2] * r[5] + r[3] * r[6]);
  }
  vol0 /= 3.0;
  const cplx nj(0.0, -1.0), np(0.0, 1.0);
#pragma omp parallel for schedule(static)
  for (int64_t iz = 0; iz < nq; ++iz) {
    int64_t iy = iz % npix;
    cplx acc = 0.0;
    bool zero = false;
    for (int64_t t = 0; t < ntri && !zero; ++t) {
      const double* r = rec + 8 * t;
      std::vector<cplx> mq(3);
      rotate_q(rot, qx[iy], qy[iy], qz_at(qz, iz), mq.data());
      double q2 = std::norm(mq[0]) + std::norm(mq[1]) + std::norm(mq[2]);
      if (q2 == 0.0) {
        zero = true;
        break;
      }
      cplx qn = mq[0] * r[1] + mq[1] * r[2] + mq[2] * r[3];
      cplx qt = mq[0] * r[4] + mq[1] * r[5] + mq[2] * r[6];
      acc += (nj * qn * r[0] * std::exp(np * qt) / q2);
    }
    store(ff, iz, zero ? cplx(vol0, 0.0) : acc);
  }
}

// Record adjoint of the approximate form factor (no reference counterpart: the reference has
// only finite-difference gradients, src/analyzer/hipgisaxs_fit_lmvm.cpp:190-224). For
// L = Re sum_q W(q) F(q), with F as in hgxo_ff_approx and A_t = s_t n_t:
//   dL/dA_t = Re sum_q W (-i / |q|^2) mq exp(i mq.c_t)
//   dL/dc_t = Re sum_q W (mq.A_t) / |q|^2 mq exp(i mq.c_t)
// and at q = 0 (F = V = sum_t A_t.c_t / 3): dL/dA_t += Re W c_t / 3, dL/dc_t += Re W A_t / 3.
// grad[t] = (dL/dA_t, dL/dc_t). Pinned by central finite differences of hgxo_ff_approx.
void hgxo_ff_approx_adjoint(int64_t ntri, const double* rec, int64_t npix, int64_t nq, const double* qx,
                            const double* qy, const double* qz, const double* rot, const double* W,
                            double* grad) {
  std::vector<cplx> mq(3 * nq);
  std::vector<double> inv_q2(nq);
  for (int64_t iz = 0; iz < nq; ++iz) {
    int64_t iy = iz % npix;
    rotate_q(rot, qx[iy], qy[iy], qz_at(qz, iz), &mq[3 * iz]);
    double q2 = std::norm(mq[3 * iz]) + std::norm(mq[3 * iz + 1]) + std::norm(mq[3 * iz + 2]);
    inv_q2[iz] = q2 == 0.0 ? 0.0 : 1.0 / q2;
  }
  const cplx nj(0.0, -1.0), np(0.0, 1.0);
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < ntri; ++t) {
    const double* r = rec + 8 * t;
    const double A[3] = {r[0] * r[1], r[0] * r[2], r[0] * r[3]};
    const double c[3] = {r[4], r[5], r[6]};
    double gA[3] = {0, 0, 0}, gc[3] = {0, 0, 0};
    for (int64_t iz = 0; iz < nq; ++iz) {
      const cplx w(W[2 * iz], W[2 * iz + 1]);
      const cplx* m = &mq[3 * iz];
      if (inv_q2[iz] == 0.0) {  // q = 0: F = V
        for (int k = 0; k < 3; ++k) {
          gA[k] += w.real() * c[k] / 3.0;
          gc[k] += w.real() * A[k] / 3.0;
        }
        continue;
      }
      cplx e = std::exp(np * (m[0] * c[0] + m[1] * c[1] + m[2] * c[2]));
      cplx qa = m[0] * A[0] + m[1] * A[1] + m[2] * A[2];
      cplx wa = w * nj * inv_q2[iz] * e;   // multiplies mq for dL/dA
      cplx wc = w * qa * inv_q2[iz] * e;   // multiplies mq for dL/dc
      for (int k = 0; k < 3; ++k) {
        gA[k] += (wa * m[k]).real();
        gc[k] += (wc * m[k]).real();
      }
    }
    for (int k = 0; k < 3; ++k) {
      grad[6 * t + k] = gA[k];
      grad[6 * t + 3 + k] = gc[k];
    }
  }
}

// ---------------------------------------------------------------------------------------------
// a8: FormFactorTriangle / compute_exact_triangle (src/ff/cpu/ff_tri_cpu.cpp:46-195).
// The reference evaluates the exact (Green's theorem) triangle integral with phase exp(-i q.v);
// sign = -1 reproduces that; sign = +1 evaluates it at -q, i.e. the exp(+i q.r) convention of the
// approximate and analytic form factors (F_+(q) = F_-(-q)).
// thresholds: tiny (the reference CPU uses DBL_MIN, its GPU 1e-13 -- ff_tri_gpu.cu:80,100).
// Deviations: case 3 carries +i (see below); at |q|^2 == 0 exactly the volume
// sum_t area_t (n_t.v1_t) / 3 is returned.
// Pinned by: a 12-triangle box reproduces the analytic box to round-off; F(0) = V; brute force.
static cplx exact_triangle(const cplx mq[3], double q_sqr, const double* tri, double tiny) {
  const cplx unitc(0.0, 1.0), n_unitc(0.0, -1.0);
  double v[3][3] = {{tri[0], tri[1], tri[2]}, {tri[3], tri[4], tri[5]}, {tri[6], tri[7], tri[8]}};
  double e[3][3];
  for (int i = 0; i < 3; ++i) {
    e[0][i] = v[1][i] - v[0][i];
    e[1][i] = v[2][i] - v[1][i];
    e[2][i] = v[0][i] - v[2][i];
  }
  double nt[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                  e[0][0] * e[1][1] - e[0][1] * e[1][0]};
  double ntn = std::sqrt(nt[0] * nt[0] + nt[1] * nt[1] + nt[2] * nt[2]);
  double area = 0.5 * ntn;
  for (int i = 0; i < 3; ++i) nt[i] /= ntn;
  cplx q_nt = 0.0;
  for (int i = 0; i < 3; ++i) q_nt += mq[i] * nt[i];
  double proj_tq = q_sqr - std::norm(q_nt);
  cplx ff = 0.0;
  if (std::abs(proj_tq) < tiny) {  // case 1 (:87-91)
    cplx q_v = 0.0;
    for (int i = 0; i < 3; ++i) q_v += mq[i] * v[0][i];
    return unitc * q_nt * area / q_sqr * std::exp(n_unitc * q_v);
  }
  for (int k = 0; k < 3; ++k) {
    double ne[3] = {e[k][1] * nt[2] - e[k][2] * nt[1], e[k][2] * nt[0] - e[k][0] * nt[2],
                    e[k][0] * nt[1] - e[k][1] * nt[0]};
    double nen = std::sqrt(ne[0] * ne[0] + ne[1] * ne[1] + ne[2] * ne[2]);
    for (int i = 0; i < 3; ++i) ne[i] /= nen;
    cplx q_ne = 0.0;
    for (int i = 0; i < 3; ++i) q_ne += mq[i] * ne[i];
    double proj_eq = proj_tq - std::norm(q_ne);
    double elen = std::sqrt(e[k][0] * e[k][0] + e[k][1] * e[k][1] + e[k][2] * e[k][2]);
    cplx q_v = 0.0;
    for (int i = 0; i < 3; ++i) q_v += mq[i] * v[k][i];
    if (std::abs(proj_eq) < tiny) {  // case 2 (:107-114)
      double f0 = elen / (q_sqr * proj_tq);
      cplx c0 = -q_nt * q_ne;
      ff += f0 * c0 * std::exp(n_unitc * q_v);
    } else {  // case 3 (:115-149)
      double f0 = q_sqr * proj_tq * proj_eq;
      double nv[3] = {e[k][0] / elen, e[k][1] / elen, e[k][2] / elen};
      cplx q_nv = 0.0;
      for (int i = 0; i < 3; ++i) q_nv += mq[i] * nv[i];
      // The reference multiplies case 3 by -i (:132, :146), the opposite sign of its cases 1-2
      // and of the volume integral (a closed box then comes out as -F); +i is used here, pinned
      // by the analytic box.
      ff += unitc * q_nt * q_ne * q_nv * std::exp(n_unitc * q_v) / f0;
      int kp = (k + 1) % 3;
      cplx q_v2 = 0.0;
      for (int i = 0; i < 3; ++i) q_v2 += mq[i] * v[kp][i];
      cplx q_nv2 = 0.0;
      for (int i = 0; i < 3; ++i) q_nv2 -= mq[i] * nv[i];
      ff += unitc * q_nt * q_ne * q_nv2 * std::exp(n_unitc * q_v2) / f0;
    }
  }
  return ff;
}

void hgxo_ff_exact(int64_t ntri, const double* tri, int64_t npix, int64_t nq, const double* qx,
                   const double* qy, const double* qz, const double* rot, int sign, double tiny,
                   double* ff) {
  double vol0 = 0.0;
  for (int64_t t = 0; t < ntri; ++t) {
    const double* v = tri + 9 * t;
    double a[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
    double b[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    vol0 += (c[0] * v[0] + c[1] * v[1] + c[2] * v[2]) / 6.0;  // area * (n.v1) / 3
  }
  if (tiny <= 0) tiny = kTiny;
  double s = sign >= 0 ? -1.0 : 1.0;  // +1 convention: evaluate the reference formula at -q
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t iz = 0; iz < nq; ++iz) {
    int64_t iy = iz % npix;
    cplx mq[3];
    rotate_q(rot, qx[iy], qy[iy], qz_at(qz, iz), mq);
    for (int j = 0; j < 3; ++j) mq[j] *= s;
    double q2 = std::norm(mq[0]) + std::norm(mq[1]) + std::norm(mq[2]);
    if (q2 == 0.0) { store(ff, iz, cplx(vol0, 0.0)); continue; }
    cplx acc = 0.0;
    for (int64_t t = 0; t < ntri; ++t) acc += exact_triangle(mq, q2, tri + 9 * t, tiny);
    store(ff, iz, acc);
  }
}

// ---------------------------------------------------------------------------------------------
// a10: AnalyticFormFactor::param_distribution (src/ff/ff_ana.cpp:235-284).
// flags & 1 (reference form): nodes by the reference's float-accumulating do-while (:249-254),
// which can yield nvalues +- 1 nodes; weights stat 0 (none) -> single weight 1, 1 (uniform) -> 1
// each, 2 (gaussian) -> unnormalised pdf exp(-(x-mean)^2 / 2sd^2) / (sqrt(2 pi) sd) (:271-272); a
// non-finite mean is replaced by the midpoint of the nodes (:266-268).
// flags 0 (default): exactly nvalues nodes pmin + k (pmax - pmin) / (nvalues - 1) and the same
// weights normalised to sum 1 (stat 1: 1/n each; stat 0 keeps the single weight 1).
// Returns the node count (or -1), writing at most cap nodes.
// Pinned by: tests (node grids against numpy linspace, the do-while count on known cases; pdf against
// scipy.stats; normalisation).
int hgxo_param_distribution(double pmin, double pmax, int nvalues, int stat, double mean, double sd, int flags,
                            double* vals, double* wts, int cap) {
  if (nvalues < 1) return -1;
  std::vector<double> dim, w;
  double pmx = pmax;
  if (pmx < pmin) pmx = pmin;
  const bool reference = (flags & 1) != 0;
  if (nvalues > 1 && reference) {
    double step = std::fabs(pmx - pmin) / (nvalues - 1);
    double curr = pmin;
    do {
      dim.push_back(curr);
      curr += step;
    } while (curr < pmax);
  } else if (nvalues > 1) {
    for (int k = 0; k < nvalues; ++k) dim.push_back(pmin + (pmx - pmin) * k / (nvalues - 1));
  } else {
    dim.push_back(pmin);
  }
  if (stat == 0) {
    w.push_back(1.0);
  } else if (stat == 1) {
    w.assign(dim.size(), reference ? 1.0 : 1.0 / dim.size());
  } else if (stat == 2) {
    double m = mean;
    if (!std::isfinite(m)) m = (dim[0] + dim[dim.size() - 1]) / 2;
    for (size_t i = 0; i < dim.size(); ++i)
      w.push_back(std::exp(-1.0 * std::pow((dim[i] - m), 2) / (2 * std::pow(sd, 2))) / (std::sqrt(2 * kPi) * sd));
    if (!reference) {
      double sum = 0.0;
      for (double x : w) sum += x;
      if (!(sum > 0)) return -1;
      for (double& x : w) x /= sum;
    }
  } else {
    return -1;
  }
  int n = (int)dim.size();
  for (int i = 0; i < n && i < cap; ++i) {
    vals[i] = dim[i];
    wts[i] = i < (int)w.size() ? w[i] : 0.0;
  }
  return n;
}

// sinc of numeric_utils.hpp:54-64 (free hig::sinc: Taylor branch for |x| <= 1e-14).
static inline cplx sinc_free(cplx x) {
  if (std::abs(x) > 1.0e-14) return std::sin(x) / x;
  return (1. - x * x / 6. + x * x * x * x / 120.);
}
// AnalyticFormFactor::sinc member (ff_ana.cpp:224-229), used by prism6.
static inline cplx sinc_member(cplx v) {
  if (std::fabs(v.real()) <= 1e-14 && std::fabs(v.imag()) <= 1e-14) return cplx(1.0, 0.0);
  return std::sin(v) / v;
}
// AnalyticFormFactor::fq_inv (ff_ana.cpp:190-196).
static inline cplx fq_inv(cplx value, double y) {
  const cplx unitc(0, 1.0);
  cplx temp = 2.0 * std::exp(unitc * value * y / 2.0) * std::sin(value * y / 2.0) / value;
  if (std::abs(temp) <= 1e-14) temp = y;
  return temp;
}
// exp(i * t) for complex t, written as the reference does: exp(complex(-Im t, Re t)).
static inline cplx expi(cplx t) { return std::exp(cplx(-t.imag(), t.real())); }

static inline cplx trans_phase(const cplx mq[3], const double* tv) {
  if (!tv) return 1.0;
  return expi(mq[0] * tv[0] + mq[1] * tv[1] + mq[2] * tv[2]);
}

// a11: FormFactorSphere / compute_sphere (src/ff/ff_ana_sphere.cpp:33-118).
// F = sum_r w_r 4 pi R^3 (sin qR - qR cos qR)/(qR)^3 exp(i qz R) * exp(i q.t), q = sqrt(q.q)
// (complex, not Hermitian); F = 0 for |q| < TINY_.
// Pinned by: scipy spherical_jn: 4 pi R^3 j1(qR)/(qR) (tests/test_oracle_analytic.py).
void hgxo_ff_sphere(int nr, const double* r, const double* wr, int64_t npix, int64_t nq, const double* qx,
                    const double* qy, const double* qz, const double* rot, const double* transvec, double* ff) {
  const cplx I(0.0, 1.0);
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < nq; ++z) {
    int64_t y = z % npix;
    cplx mq[3];
    rotate_q(rot, qx[y], qy[y], qz_at(qz, z), mq);
    cplx qval = std::sqrt(mq[0] * mq[0] + mq[1] * mq[1] + mq[2] * mq[2]);
    cplx acc = 0.0;
    if (!(std::abs(qval) < kTiny)) {
      for (int i = 0; i < nr; ++i) {
        double R = r[i];
        cplx qR = qval * R;
        cplx qR3 = qR * qR * qR;
        cplx t1 = std::conj(qR3) / std::norm(qR3);
        cplx c0 = (std::sin(qR) - qR * std::cos(qR)) * t1;
        cplx c1 = std::exp(I * mq[2] * R);
        acc += wr[i] * (4 * kPi * R * R * R) * c0 * c1;
      }
    }
    store(ff, z, acc * trans_phase(mq, transvec));
  }
}

// a12: FormFactorCylinder / compute_cylinder (src/ff/ff_ana_cylinder.cpp:34-126).
// F = 2 pi R^2 H J1(q_par R)/(q_par R) sinc(qz H/2) exp(i qz H/2); q_par = sqrt(mqx^2 + mqy^2).
// J1 follows cbessj (src/numerics/numeric_utils.cpp:69-71): J1 of the real part only; the
// reference's division uses conj(t)/|t|^2 with t complex, kept. The node sum carries weights
// w_r w_h (the reference CPU sums unweighted, :107-111: pass unit weights to reproduce it).
// Pinned by: scipy j1 closed form on real q.
void hgxo_ff_cylinder(int nr, const double* r, const double* wr, int nh, const double* h, const double* wh,
                      int64_t npix, int64_t nq, const double* qx, const double* qy, const double* qz,
                      const double* rot, const double* transvec, double* ff) {
  const cplx I(0.0, 1.0);
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < nq; ++z) {
    int64_t y = z % npix;
    cplx mq[3];
    rotate_q(rot, qx[y], qy[y], qz_at(qz, z), mq);
    cplx qpar = std::sqrt(mq[0] * mq[0] + mq[1] * mq[1]);
    cplx acc = 0.0;
    for (int i = 0; i < nr; ++i)
      for (int k = 0; k < nh; ++k) {
        double R = r[i], H = h[k];
        double vol = 2. * kPi * R * R * H;
        cplx sv = sinc_free(0.5 * mq[2] * H);
        cplx ev = std::exp(I * 0.5 * mq[2] * H);
        cplx t1 = qpar * R;
        cplx bess = 0.0;
        if (std::norm(t1) > kCuTiny) bess = cplx(::j1(t1.real()), 0.0) * std::conj(t1) / std::norm(t1);
        acc += wr[i] * wh[k] * (vol * sv * ev * bess);
      }
    store(ff, z, acc * trans_phase(mq, transvec));
  }
}

// a13: FormFactorBox / compute_box (src/ff/ff_ana_box.cpp:35-145) and FormFactorCube
// (src/ff/ff_ana_cube.cpp:35-130, a box with lx = ly = lz = edge).
// F = LWH sinc(qx L/2) sinc(qy W/2) sinc(qz H/2) exp(i qz H/2), weighted wx wy wz.
// Pinned by: closed-form product of sincs; the exact-triangle mesh of the same box.
void hgxo_ff_box(int nx, const double* x, const double* wx, int ny, const double* yv, const double* wy, int nz,
                 const double* zv, const double* wz, int64_t npix, int64_t nq, const double* qx, const double* qy,
                 const double* qz, const double* rot, const double* transvec, double* ff) {
  const cplx I(0.0, 1.0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nq; ++i) {
    int64_t j = i % npix;
    cplx mq[3];
    rotate_q(rot, qx[j], qy[j], qz_at(qz, i), mq);
    cplx acc = 0.0;
    for (int iz = 0; iz < nz; ++iz)
      for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
          double L = x[ix], W = yv[iy], H = zv[iz];
          double wght = wx[ix] * wy[iy] * wz[iz];
          cplx ev = std::exp(I * 0.5 * mq[2] * H);
          cplx sv = sinc_free(0.5 * mq[0] * L) * sinc_free(0.5 * mq[1] * W) * sinc_free(0.5 * mq[2] * H);
          acc += (L * W * H) * ev * sv * wght;
        }
    store(ff, i, acc * trans_phase(mq, transvec));
  }
}

// a13: FormFactorCube / compute_cube (src/ff/ff_ana_cube.cpp:35-42, :103-113): one edge node list,
// F = sum_i w_i L_i^3 sinc(qx L/2) sinc(qy L/2) sinc(qz L/2) exp(i qz L/2).
// Pinned by: the box closed form with L = W = H.
void hgxo_ff_cube(int nx, const double* x, const double* wx, int64_t npix, int64_t nq, const double* qx,
                  const double* qy, const double* qz, const double* rot, const double* transvec, double* ff) {
  const cplx I(0.0, 1.0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nq; ++i) {
    int64_t j = i % npix;
    cplx mq[3];
    rotate_q(rot, qx[j], qy[j], qz_at(qz, i), mq);
    cplx acc = 0.0;
    for (int ix = 0; ix < nx; ++ix) {
      double L = x[ix];
      cplx ev = std::exp(I * 0.5 * mq[2] * L);
      cplx sv = sinc_free(0.5 * mq[0] * L) * sinc_free(0.5 * mq[1] * L) * sinc_free(0.5 * mq[2] * L);
      acc += (L * L * L) * ev * sv * wx[ix];
    }
    store(ff, i, acc * trans_phase(mq, transvec));
  }
}

// a14: compute_prism, triangular prism (src/ff/ff_ana_prism.cpp:34-122; CPU formula :87-113).
// Node sum carries weights w_l w_h (the reference CPU ignores the weights: pass unit weights).
// Pinned by: the exact-triangle mesh of the same prism (tests).
void hgxo_ff_prism3(int nl, const double* l, const double* wl, int nh, const double* h, const double* wh,
                    double tau, double eta, int64_t npix, int64_t nq, const double* qx, const double* qy,
                    const double* qz, const double* rot, const double* transvec, double* ff) {
  const double sqrt3 = std::sqrt(3.0);
  const cplx unitc(0, 1.0);
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < nq; ++z) {
    int64_t y = z % npix;
    cplx mq[3];
    rotate_q(rot, qx[y], qy[y], qz_at(qz, z), mq);
    cplx mqx = mq[0], mqy = mq[1], mqz = mq[2];
    cplx temp1 = mqx * (mqx * mqx - 3.0 * mqy * mqy);
    cplx temp2 = mqz + std::tan(tau) * (mqx * std::sin(eta) + mqy * std::cos(eta));
    cplx acc = 0.0;
    for (int il = 0; il < nl; ++il)
      for (int ih = 0; ih < nh; ++ih) {
        cplx temp4 = mqx * std::exp(unitc * mqy * l[il] * sqrt3);
        cplx temp5 = mqx * std::cos(mqx * l[il]);
        cplx temp6 = unitc * sqrt3 * mqy * std::sin(mqx * l[il]);
        cplx temp7 = fq_inv(temp2, h[ih]);
        cplx temp8 = (temp4 - temp5 - temp6) * temp7;
        cplx temp3 = mqy * l[il] / sqrt3;
        cplx temp9 = 2.0 * sqrt3 * std::exp(cplx(temp3.imag(), -temp3.real()));
        acc += wl[il] * wh[ih] * ((temp9 / temp1) * temp8);
      }
    store(ff, z, acc * trans_phase(mq, transvec));
  }
}

// a14: compute_prism6, hexagonal prism (src/ff/ff_ana_prism6.cpp:34-109; CPU formula :81-101).
// Weighted node sum (the reference CPU ignores the weights: pass unit weights).
// Pinned by: the exact-triangle mesh of the same hexagonal prism (tests).
void hgxo_ff_prism6(int nl, const double* l, const double* wl, int nh, const double* h, const double* wh,
                    double tau, double eta, int64_t npix, int64_t nq, const double* qx, const double* qy,
                    const double* qz, const double* rot, const double* transvec, double* ff) {
  const double sqrt3 = std::sqrt(3.0);
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < nq; ++z) {
    int64_t y = z % npix;
    cplx mq[3];
    rotate_q(rot, qx[y], qy[y], qz_at(qz, z), mq);
    cplx qm = std::tan(tau) * (mq[0] * std::sin(eta) + mq[1] * std::cos(eta));
    cplx temp1 = (4.0 * sqrt3) / (3.0 * mq[1] * mq[1] - mq[0] * mq[0]);
    cplx acc = 0.0;
    for (int ih = 0; ih < nh; ++ih)
      for (int il = 0; il < nl; ++il) {
        cplx rmqx = l[il] * mq[0] / sqrt3;
        cplx rmqy = l[il] * mq[1];
        cplx temp2 = rmqy * rmqy * sinc_member(rmqx) * sinc_member(rmqy);
        cplx temp3 = std::cos(2.0 * rmqx);
        cplx temp4 = std::cos(rmqy) * std::cos(rmqx);
        cplx temp5 = temp1 * (temp2 + temp3 - temp4);
        cplx temp6 = fq_inv(mq[2] + qm, h[ih]);
        acc += wl[il] * wh[ih] * (temp5 * temp6);
      }
    store(ff, z, acc * trans_phase(mq, transvec));
  }
}

// a15: FormFactorPyramid / compute_pyramid (src/ff/ff_ana_pyramid.cpp:35-156), truncated square
// pyramid; returns 0 when 2H/L or 2H/W >= tan(angle) (untruncated) and when |qx qy| < 1e-20.
// Weighted 4-D node sum (:133-143). angle in degrees.
// compat = 1 evaluates the reference's k1..k4 (:47-60), whose exponentials exp(-/+ i q_j H) carry
// the opposite sign of the published truncated-pyramid form factor while the +-i factors of k2, k4
// keep theirs, so it matches neither the pyramid nor its mirror image. compat = 0 (default) uses
// exp(+/- i q_j H), the volume integral over the pyramid with base z = 0 and top z = H.
// Pinned by: compat = 0 against Gauss-Legendre slice integration of the pyramid and against the
// exact-triangle mesh (tests/test_oracle_analytic.py); compat = 1 only against a literal
// transcription of the reference formula (parity unpinned by any external truth).
static cplx pyramid_one(cplx qx, cplx qy, cplx qz, double length, double width, double height, double angle,
                        int compat) {
  double a = angle * kPi / 180.;
  double tan_a = std::tan(a);
  if ((2 * height / length >= tan_a) || (2 * height / width >= tan_a)) return 0.0;
  cplx tmp = qx * qy;
  if (std::abs(tmp) < 1.0E-20) return 0.0;
  const cplx P_J(0, 1), N_J(0, -1);
  const cplx E1 = compat ? N_J : P_J, E2 = compat ? P_J : N_J;
  cplx q1 = 0.5 * ((qx - qy) / tan_a + qz);
  cplx q2 = 0.5 * ((qx - qy) / tan_a - qz);
  cplx q3 = 0.5 * ((qx + qy) / tan_a + qz);
  cplx q4 = 0.5 * ((qx + qy) / tan_a - qz);
  cplx a1 = sinc_free(q1 * height) * std::exp(E1 * q1 * height);
  cplx a2 = sinc_free(q2 * height) * std::exp(E2 * q2 * height);
  cplx a3 = sinc_free(q3 * height) * std::exp(E1 * q3 * height);
  cplx a4 = sinc_free(q4 * height) * std::exp(E2 * q4 * height);
  cplx k1 = a1 + a2;
  cplx k2 = a1 * N_J + a2 * P_J;
  cplx k3 = a3 + a4;
  cplx k4 = a3 * N_J + a4 * P_J;
  cplx t1 = k1 * std::cos(0.5 * (qx * length - qy * width));
  cplx t2 = k2 * std::sin(0.5 * (qx * length - qy * width));
  cplx t3 = k3 * std::cos(0.5 * (qx * length + qy * width));
  cplx t4 = k4 * std::sin(0.5 * (qx * length + qy * width));
  return ((height / tmp) * (t1 + t2 - t3 - t4));
}

void hgxo_ff_pyramid(int nx, const double* x, const double* wx, int ny, const double* yv, const double* wy, int nh,
                     const double* h, const double* wh, int nb, const double* b, const double* wb, int compat,
                     int64_t npix, int64_t nq, const double* qx, const double* qy, const double* qz, const double* rot,
                     const double* transvec, double* ff) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nq; ++i) {
    int64_t j = i % npix;
    cplx mq[3];
    rotate_q(rot, qx[j], qy[j], qz_at(qz, i), mq);
    cplx acc = 0.0;
    for (int ix = 0; ix < nx; ++ix)
      for (int iy = 0; iy < ny; ++iy)
        for (int ih = 0; ih < nh; ++ih)
          for (int ib = 0; ib < nb; ++ib) {
            double prob = wx[ix] * wy[iy] * wh[ih] * wb[ib];
            acc += pyramid_one(mq[0], mq[1], mq[2], x[ix], yv[iy], h[ih], b[ib], compat) * prob;
          }
    store(ff, i, acc * trans_phase(mq, transvec));
  }
}

// ---------------------------------------------------------------------------------------------
// a16: StructureFactor::compute_structure_factor, lattice branch (src/sf/sf.cpp:165-245).
// S = exp(i q.center) prod_k pow(e_k, (1-N_k)/2) (1 - e_k^N_k)/(1 - e_k), e_k = exp(i l_k.q),
// with the x -> 0 limit N_k when |1 - e_k| <= eps per component (:197-232). pow() is the
// principal branch, as std::pow does. la, lb, lc are the lattice vectors already multiplied by
// the grain scaling (:169-171).
// flags bit 0: reproduce the reference's b-term typo lb[2]*lb[2] in place of lb[2]*mq[2] (:212).
// Addition (not in the reference): Debye-Waller damping exp(-sigma^2 |q|^2 / 2),
// |q|^2 = sum_j |mq_j|^2; dw_sigma = 0 disables it.
// Pinned by: the direct lattice sum over N_a N_b N_c cells and the Dirichlet identity
// (1-e^{iNx})/(1-e^{ix}) e^{ix(1-N)/2} = sin(Nx/2)/sin(x/2) (tests/test_oracle_sf.py).
void hgxo_sf_lattice(const double* la, const double* lb, const double* lc, const double* repet,
                     const double* center, double dw_sigma, int flags, int64_t npix, int64_t nq,
                     const double* qx, const double* qy, const double* qz, const double* rot, double* sf) {
  double rep[3] = {repet[0] < 1 ? 1 : repet[0], repet[1] < 1 ? 1 : repet[1], repet[2] < 1 ? 1 : repet[2]};
  const cplx unit_c(1, 0), unit_ci(0, 1);
  const double mach_eps = std::numeric_limits<double>::epsilon();
  const bool typo = (flags & 1) != 0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nq; ++i) {
    int64_t j = i % npix;
    cplx mq[3];
    rotate_q(rot, qx[j], qy[j], qz_at(qz, i), mq);
    cplx sa, sb, sc;
    cplx e_iqa = std::exp(unit_ci * (la[0] * mq[0] + la[1] * mq[1] + la[2] * mq[2]));
    cplx Xa_0 = unit_c - std::pow(e_iqa, rep[0]);
    cplx Ya_0 = unit_c - e_iqa;
    if (std::fabs(Ya_0.imag()) > mach_eps || std::fabs(Ya_0.real()) > mach_eps) sa = Xa_0 / Ya_0;
    else sa = rep[0];
    sa = std::pow(e_iqa, (1.0 - rep[0]) / 2.0) * sa;

    cplx bz = typo ? cplx(lb[2] * lb[2], 0.0) : lb[2] * mq[2];
    cplx iqb = unit_ci * (lb[0] * mq[0] + lb[1] * mq[1] + bz);
    cplx iqNb = rep[1] * iqb;
    cplx e_iqb = std::exp(iqb);
    cplx Xb_0 = unit_c - std::exp(iqNb);
    cplx Yb_0 = unit_c - std::exp(iqb);
    if (std::fabs(Yb_0.imag()) > mach_eps || std::fabs(Yb_0.real()) > mach_eps) sb = Xb_0 / Yb_0;
    else sb = rep[1];
    sb = std::pow(e_iqb, (1.0 - rep[1]) / 2.0) * sb;

    cplx e_iqc = std::exp(unit_ci * (lc[0] * mq[0] + lc[1] * mq[1] + lc[2] * mq[2]));
    cplx Xc_0 = unit_c - std::pow(e_iqc, rep[2]);
    cplx Yc_0 = unit_c - e_iqc;
    if (std::fabs(Yc_0.imag()) > mach_eps || std::fabs(Yc_0.real()) > mach_eps) sc = Xc_0 / Yc_0;
    else sc = rep[2];
    sc = std::pow(e_iqc, (1.0 - rep[2]) / 2.0) * sc;

    cplx temp3 = center[0] * mq[0] + center[1] * mq[1] + center[2] * mq[2];
    temp3 = std::exp(cplx(-temp3.imag(), temp3.real()));
    cplx s = temp3 * sa * sb * sc;
    if (dw_sigma != 0.0) {
      double q2 = std::norm(mq[0]) + std::norm(mq[1]) + std::norm(mq[2]);
      s *= std::exp(-0.5 * dw_sigma * dw_sigma * q2);
    }
    store(sf, i, s);
  }
}

// ---------------------------------------------------------------------------------------------
// a17: DWBA combination (src/sim/hipgisaxs_main.cpp:1084-1100):
// I[p] += |w * sum_{k=0..3} fc_k sf_k ff_k|^2 over the 4 branches of pixel p.
// nbranch = 1 gives the SAXS/Born form |w sf ff|^2 (:1102-1105; accumulated here).
// Pinned by: delta = beta = 0 (fc = [1,0,0,0]) reduces to the Born |F|^2; linearity in fc.
void hgxo_dwba_intensity(int64_t npix, int nbranch, const double* fc, const double* sf, const double* ff,
                         double weight, double* I) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < npix; ++i) {
    cplx t = 0.0;
    for (int k = 0; k < nbranch; ++k) {
      int64_t idx = i + k * npix;
      cplx f = fc ? qz_at(fc, idx) : cplx(1.0, 0.0);
      cplx s = sf ? qz_at(sf, idx) : cplx(1.0, 0.0);
      t += f * s * qz_at(ff, idx);
    }
    t *= weight;
    I[i] += t.real() * t.real() + t.imag() * t.imag();
  }
}

// Mesh helpers used by tests: exact (9-real) triangles -> approximate 8-real records
// (objectshape_reader.cpp:271-300: area from |a||b| sin(theta)/2, unit normal a x b, centroid).
// Degenerate triangles are skipped as the reader does (:293). Returns the record count.
int64_t hgxo_tri9_to_rec8(int64_t ntri, const double* tri, double* rec) {
  int64_t n = 0;
  for (int64_t t = 0; t < ntri; ++t) {
    const double* v = tri + 9 * t;
    double a[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
    double b[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    double na = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    double nb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double nc = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    if (nc == 0 || na == 0 || nb == 0) continue;
    double ct = dot / (na * nb);
    double sint = std::sqrt(1 - ct * ct);
    double* r = rec + 8 * n;
    r[0] = na * nb * sint / 2;
    r[1] = c[0] / nc; r[2] = c[1] / nc; r[3] = c[2] / nc;
    r[4] = (v[0] + v[3] + v[6]) / 3.0;
    r[5] = (v[1] + v[4] + v[7]) / 3.0;
    r[6] = (v[2] + v[5] + v[8]) / 3.0;
    r[7] = 0.0;
    ++n;
  }
  return n;
}

}  // extern "C"
