/* hgx_hig.h -- C ABI of the .hig simulation driver (libhgx_hig.so, host code on the hgx C ABI).
 *
 * Replaces the reference's HipGISAXS driver object (include/sim/hipgisaxs_main.hpp, used by
 * src/sim/hipgisaxs_sim.cpp:37-70): construct_input, init, compute_gisaxs / run_all_gisaxs and
 * save_gisaxs. Images are host arrays of nrow * ncol doubles, row-major, row 0 holding the highest
 * q_z (src/model/qgrid.cpp:75). Every function returns HGX_OK (0) or a negative HGX_ERR_* code
 * (include/hgx.h) and never exits; hgx_hig_last_error returns the message of the last failure.
 */
#ifndef HGX_HIG_H_
#define HGX_HIG_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hgx_hig hgx_hig;

/* flags of hgx_hig_options */
#define HGX_HIG_NO_FUSE 1           /* per-grain form factor, structure factor and DWBA calls only */
#define HGX_HIG_SF_COMPAT_LB2 2     /* the reference CPU structure factor's lb[2]*lb[2] term */
#define HGX_HIG_SMEAR_COMPAT 4      /* the reference's smearing window (HGX_SMEAR_COMPAT) */
#define HGX_HIG_NO_SMEAR 8          /* skip the smearing whatever the input's sigma */
#define HGX_HIG_MESH_NO_TRANSLATION 16 /* ignore unit-cell locations of mesh shapes, as the reference
                                          does (src/ff/ff.cpp:44) */
#define HGX_HIG_PROFILE 32          /* time every kernel of a run with HIP events (hgx_hig_kernel_ms) */
#define HGX_HIG_FIT_FD_GRADIENT 64  /* LMVM gradients by the reference's central differences only */
#define HGX_HIG_FIT_SEQUENTIAL 128  /* particle swarm: one synchronisation per particle, as the reference */
#define HGX_HIG_TRUNCATED_SPHERE 256 /* spheres with a height parameter cut as the reference's GPU path cuts
                                        them (HGX_SHAPE_TRUNCATED_SPHERE); by default heights are ignored,
                                        as its CPU path does */

/* A driver on HIP device `device`. Reading inputs, hgx_hig_info, hgx_hig_config and hgx_hig_grains
 * run on the host; the device context (one hgx context) is created by the first call that needs it. */
int hgx_hig_create(int device, hgx_hig** out);
int hgx_hig_destroy(hgx_hig* h);
/* HipGISAXS::construct_input (src/sim/hipgisaxs_main.cpp, HiGInput::construct_input_config
 * src/config/hig_input.cpp:62-106): reads a .hig file, or .hig text with base_dir for relative mesh
 * file names. Resets the orientation sampler to the seed of hgx_hig_options. */
int hgx_hig_load(hgx_hig* h, const char* path);
int hgx_hig_load_text(hgx_hig* h, const char* text, const char* base_dir);
/* precision: HGX_PREC_FP64 (default) or HGX_PREC_FP32C; flags: HGX_HIG_*; seed of the glibc rand()
 * sequence the orientation statistics "range" and "random" draw from (1, the default, is the
 * sequence of a reference process, which never seeds it on that path). Resets the sampler. */
int hgx_hig_options(hgx_hig* h, int precision, int flags, unsigned seed);
/* HipGISAXS::init (src/sim/hipgisaxs_main.cpp:106-216): k0 from the photon energy and the q-grid of
 * the output region (q-space) at the smallest incidence angle. Called by the runs when needed. */
int hgx_hig_init(hgx_hig* h);
/* Image size, number of (alpha_i, phi, tilt) configurations run_all_gisaxs scans, k0 (1/nm).
 * Host only (no device work). */
int hgx_hig_info(hgx_hig* h, int* ncol, int* nrow, int* nconfig, double* k0);
/* Angles (degrees) of configuration `config` in run_all_gisaxs's loop order (alpha_i outermost). */
int hgx_hig_config(hgx_hig* h, int config, double* alpha_i, double* phi, double* tilt);
/* run_gisaxs for configuration `config`: image [nrow * ncol] (host). Draws grain orientations from
 * the sampler, as each reference run does. */
int hgx_hig_run(hgx_hig* h, int config, double* image);
/* compute_gisaxs: the single configuration of a non-scanning input. */
int hgx_hig_compute(hgx_hig* h, double* image);
/* The grains the next run gives structure `key` (host [cap][9] rotations applied to q, [cap][3]
 * centres, [cap] weights; any may be NULL) without advancing the sampler; *n = grain count. */
int hgx_hig_grains(hgx_hig* h, const char* key, int cap, double* rot, double* center, double* weight, int* n);
/* Work of the last run: counts[0] grains through the fused lattice kernel, [1] grains through
 * per-grain calls, [2] form-factor evaluations, [3] structure-factor evaluations; *ms its wall time. */
int hgx_hig_last_counts(hgx_hig* h, int64_t counts[4], double* ms);
/* Numeric (mesh) form-factor calls of the last run by the kernels they took (hgx_ff_numeric_path of
 * include/hgx.h): paths[0] per-pair kernels, [1] product grid on the matrix cores, [2] Ewald-sphere
 * grid on the matrix cores. No reference counterpart. */
int hgx_hig_last_paths(hgx_hig* h, int64_t paths[3]);
/* With HGX_HIG_PROFILE: summed durations (ms) and launch counts of the last run's kernels, "main" the
 * form-factor / structure-factor / lattice kernels, "aux" the rest (hgx_profile_end of include/hgx.h). */
int hgx_hig_kernel_ms(hgx_hig* h, double* main_ms, int64_t* main_launches, double* aux_ms, int64_t* aux_launches);
/* HipGISAXS::save_gisaxs (src/sim/hipgisaxs_main.cpp:2250-2260): nrow lines of ncol tab-terminated
 * values. */
int hgx_hig_save(const double* image, int ncol, int nrow, const char* path);
/* The image file of run_all_gisaxs (Image::construct_image + Image::save, src/image/image.cpp:125-226,
 * 502-512): pixels shifted to non-negative, log10 (zeros kept), normalised to [0, 1], written as an
 * uncompressed 16-bit grayscale TIFF (the reference's colour palette is not applied). Host only;
 * non-finite pixels are an error. */
int hgx_hig_save_tiff(const double* image, int ncol, int nrow, const char* path);
/* Reads a mesh file of a custom shape (host only): Wavefront .obj into HGX_MESH_EXACT9 records
 * (ObjectShapeReader, src/file/objectshape_reader.cpp:52-300; polygons are split into triangle fans
 * where the reference drops every non-triangle face) or raw .dat into HGX_MESH_APPROX8 records
 * (RawShapeReader, src/file/rawshape_reader.cpp:39-62). Writes *kind and *ntri, and the records
 * into `records` when cap >= *ntri (pass cap = 0 to query the size). */
int hgx_hig_read_mesh(const char* path, int* kind, int64_t* ntri, int64_t cap, double* records);
/* Convex hull of n points (host [n][3]) by quickhull: returns the number of triangles and writes up
 * to cap of them into tri [cap][3] (point indices, counter-clockwise seen from outside; tri may be
 * NULL to count), or HGX_ERR_INVALID_ARG for fewer than 4 points, non-finite or coplanar input. Host
 * only. No counterpart in the reference (its polyhedra come from mesh files); the generator of convex
 * test shapes. */
int64_t hgx_hig_convex_hull(int64_t n, const double* pts, int64_t cap, int* tri);
/* ---- fitting (the analyzer's objective and fit drivers on the device) ------------------------- */

/* HiGInput::update_params for one value (src/config/hig_input.cpp:2712-2832): the model value a fit
 * variable path names, e.g. shape['s']:param['radius']:p1, layer['film']:refindex:delta,
 * structure['st']:grain:lattice:a:x, instrumentation:scattering:alphai:min. Host only. */
int hgx_hig_update_param(hgx_hig* h, const char* variable, double value);
/* Number of fit parameters (fitparam blocks, ordered by key as the reference's maps order them) and
 * of algorithm blocks. Host only. */
int hgx_hig_fit_info(hgx_hig* h, int* nparams, int* nalgorithms);
/* Fit parameter i: its key (NUL-terminated, truncated to key_cap) and range4 = (min, max, step, init);
 * a step of -1 means none was given. Host only. */
int hgx_hig_fit_param(hgx_hig* h, int i, char* key, size_t key_cap, double* range4);
/* Algorithm block i: its algoname, the HGX_METRIC_* of its distancemetric (-1 when not supported) and
 * tol_reg = (tolerance, regularization). Host only. */
int hgx_hig_fit_algorithm(hgx_hig* h, int i, char* name, size_t name_cap, int* metric, double* tol_reg);
/* HipGISAXSObjectiveFunction::set_reference_data (src/analyzer/objective_func_hipgisaxs.cpp:110-170):
 * reads the reference data (ASCII rows, or an ESRF data file: EDFReader, src/file/edf_reader.cpp:24-150)
 * and mask, takes the reference region and size as the output
 * region when the sizes differ (HipGISAXS::override_qregion), and keeps both on the device. */
int hgx_hig_fit_init(hgx_hig* h);
/* HipGISAXSObjectiveFunction::operator() (:245-332): sets the n fit parameters to x, computes the
 * image on the device and its distance under `metric` (HGX_METRIC_*) to the reference data, plus
 * (reg / 2) |x - range midpoints|^2; a non-finite distance gives 1e8. Runs hgx_hig_fit_init first
 * when it has not run. */
int hgx_hig_fit_objective(hgx_hig* h, int n, const double* x, int metric, double reg, double* value);
/* The objective at x and its gradient in one pass (no counterpart in the reference, whose LMVM takes
 * central differences, src/analyzer/hipgisaxs_fit_lmvm.cpp:190-224): forward-mode derivatives
 * through the ensemble kernel (hgx_ensemble_intensity_tangent) and the distance
 * (hgx_image_distance_tangent) for every variable that moves only the model's tables, their tangents
 * taken by central differences of the host's table construction. computed[j] = 1 where grad[j] was
 * so formed, 0 for variables it does not cover (instrument and layer values, integer counts, cells
 * with numeric shapes), whose grad[j] is 0. The value is the objective (FP64 when the FP32C option is
 * off). */
int hgx_hig_fit_gradient(hgx_hig* h, int n, const double* x, int metric, double reg, double* value, double* grad,
                         int* computed);
/* Runs algorithm block `algorithm`: pso (ParticleSwarmOptimization, base flavour, generator seeded
 * with `seed`, each generation's objectives queued and read back with one synchronisation unless
 * HGX_HIG_FIT_SEQUENTIAL), bruteforce, lmvm (bounded L-BFGS, gradients from hgx_hig_fit_gradient and
 * central differences for the rest, or central differences alone with HGX_HIG_FIT_FD_GRADIENT) or
 * none_pounders (the
 * objective at the initial values). x_best [nparams] and *value: the best point and its objective;
 * *evals: objective calls. */
int hgx_hig_fit_run(hgx_hig* h, int algorithm, unsigned long long seed, double* x_best, double* value,
                    long long* evals);
/* Every objective evaluation since the last hgx_hig_fit_run or hgx_hig_fit_init, in call order: rows
 * [nrows][nparams + 1] of (x, value), at most cap_rows written (rows may be NULL to query *nrows). */
int hgx_hig_fit_history(hgx_hig* h, long long cap_rows, double* rows, long long* nrows);

const char* hgx_hig_last_error(const hgx_hig* h);

#ifdef __cplusplus
}
#endif

#endif /* HGX_HIG_H_ */
