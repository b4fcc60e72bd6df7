// Reference-style callers of the drop-in C++ classes (hipgisaxs_amd/host/hig_gpu.hpp), exported
// with C linkage so the GPU tests can drive them from pytest. Each function calls the adapter the way
// the reference's own code calls its GPU class:
//   numeric FF   -> src/ff/ff_num.cpp:261-280 (compute_approx_triangle, copy p_ff out, delete[] it)
//   analytic FF  -> src/ff/ff_ana_sphere.cpp:88-91 / ff_ana_pyramid.cpp:121-127
//   SF           -> src/sf/sf.cpp:288-309 (init, compute, get_sf into a new[] buffer)
//   Fresnel      -> src/sim/hipgisaxs_main.cpp:781-786 (propagation_coeffs)
#include <cstring>
#include <vector>

#include "hig_gpu.hpp"

using namespace hig;

static RotMatrix_t rot_of(const double* r) {
  RotMatrix_t R;
  if (r) std::memcpy(R.data_, r, 9 * sizeof(double));
  return R;
}

extern "C" {

int shim_numeric_approx(int nqy, double* qx, double* qy, int nqz, double* qz, int ntri, const double* rec8,
                        const double* rot, int precision, double* out) {
  NumericFormFactorG gff(2, 4, 4);  // the block shape FormFactor eff(2, 4, 4) passes (hipgisaxs_main.cpp:995)
  gff.precision(precision);
  std::vector<real_t> shape_def(rec8, rec8 + 8 * (size_t)ntri);
  complex_t* p_ff = nullptr;
  real_t kernel_time = 0.0;
  RotMatrix_t R = rot_of(rot);
  unsigned int n = gff.compute_approx_triangle(shape_def, p_ff, nqy, qx, qy, nqz, reinterpret_cast<complex_t*>(qz), R,
                                               kernel_time);
  if (n == 0 || !p_ff) return -1;
  for (int i = 0; i < nqz; i++) {
    out[2 * i] = p_ff[i].real();
    out[2 * i + 1] = p_ff[i].imag();
  }
  delete[] p_ff;
  return (int)n;
}

int shim_numeric_exact(int nqy, double* qx, double* qy, int nqz, double* qz, int ntri, const double* tri9,
                       const double* rot, double* out) {
  NumericFormFactorG gff;
  std::vector<triangle_t> tris(ntri);
  std::memcpy(tris.data(), tri9, sizeof(triangle_t) * ntri);
  complex_t* p_ff = nullptr;
  real_t kernel_time = 0.0;
  RotMatrix_t R = rot_of(rot);
  unsigned int n = gff.compute_exact_triangle(tris.data(), ntri, p_ff, nqy, qx, qy, nqz, reinterpret_cast<complex_t*>(qz),
                                              R, kernel_time);
  if (n == 0 || !p_ff) return -1;
  std::memcpy(out, p_ff, sizeof(complex_t) * nqz);
  delete[] p_ff;
  return (int)n;
}

// shape: 0 sphere (r), 1 pyramid (x, y, h, b); dims packed as counts + values + weights
int shim_analytic(int shape, int nqy, double* qx, double* qy, int nqz, double* qz, const int* counts,
                  const double* vals, const double* wts, const double* rot, const double* transvec, int compat,
                  double* out) {
  AnalyticFormFactorG gff;
  if (!gff.init(nqy, qx, qy, nqz, reinterpret_cast<complex_t*>(qz))) return -1;
  gff.compat(compat);
  int nd = shape == 0 ? 1 : 4;
  std::vector<real_vec_t> v(4), w(4);
  for (int d = 0, off = 0; d < nd; off += counts[d], ++d) {
    v[d].assign(vals + off, vals + off + counts[d]);
    w[d].assign(wts + off, wts + off + counts[d]);
  }
  real_vec_t tv(transvec, transvec + 3);
  complex_vec_t ff;
  RotMatrix_t R = rot_of(rot);
  bool ok = shape == 0 ? gff.compute_sphere(v[0], w[0], real_vec_t(), real_vec_t(), R, tv, ff)
                       : gff.compute_pyramid(0.0, 0.0, v[0], w[0], v[1], w[1], v[2], w[2], v[3], w[3], R, tv, ff);
  if (!ok) return -1;
  std::memcpy(out, ff.data(), sizeof(complex_t) * ff.size());
  return 0;
}

int shim_sf(int nqy, double* qx, double* qy, int nqz, double* qz, const double* abc, const double* repet,
            const double* scaling, const double* center, const double* rot, double dw_sigma, double* out) {
  StructureFactorG gsf;
  if (!gsf.init(nqy, qx, qy, nqz, reinterpret_cast<complex_t*>(qz))) return -1;
  vector3_t a(abc[0], abc[1], abc[2]), b(abc[3], abc[4], abc[5]), c(abc[6], abc[7], abc[8]);
  if (!gsf.compute("gisaxs", vector3_t(center[0], center[1], center[2]), a, b, c,
                   vector3_t(repet[0], repet[1], repet[2]), vector3_t(scaling[0], scaling[1], scaling[2]),
                   rot_of(rot), dw_sigma))
    return -1;
  complex_t* sf = new complex_t[nqz];
  gsf.get_sf(sf);
  std::memcpy(out, sf, sizeof(complex_t) * nqz);
  delete[] sf;
  return 0;
}

int shim_fresnel(int nl, const double* delta, const double* beta, const double* thickness, int order, double k0,
                 double alpha_i, int nrow, const double* alpha_f, int ncol, double* out) {
  MultiLayerG ml;
  std::vector<MultiLayerG::Layer> layers;
  for (int i = 0; i < nl; ++i) layers.push_back({delta[i], beta[i], thickness[i]});
  if (!ml.init(layers)) return -1;
  complex_vec_t fc;
  if (!ml.propagation_coeffs(fc, k0, alpha_i, order, real_vec_t(alpha_f, alpha_f + nrow), (unsigned)ncol)) return -1;
  std::memcpy(out, fc.data(), sizeof(complex_t) * fc.size());
  return 0;
}

}  // extern "C"
