/* hgx.h -- C ABI of the MI355X (gfx950) GISAXS form-factor / structure-factor / DWBA engine.
 *
 * Plain C: no torch, no STL, no HIP types in the signatures. Every compute entry point writes into
 * caller-owned DEVICE memory (a pointer from hipMalloc or any framework allocator on the context's
 * device) and enqueues its kernels on the context's stream; nothing blocks unless documented.
 * Inputs given as host pointers are small (rotations, parameter lists, layer stacks) or are
 * uploaded once (q-grids, meshes) and persist in the context.
 *
 * Conventions (from the reference, paths relative to the reference root):
 *   - q-grid: per-pixel real qx[npix], qy[npix]; complex qz[nq] with nq = nbranch * npix
 *     (nbranch = 4 for DWBA, 1 for Born/SAXS); q-point i <-> pixel i % npix, branch i / npix
 *     (src/model/qgrid.cpp:102-127). Row r of the detector holds pixels [r*ncol, (r+1)*ncol).
 *   - complex arrays are interleaved double pairs (re, im): a complex[nq] output is double[2*nq].
 *   - rot is a row-major 3x3 matrix applied to q: mq = rot * q (include/numerics/matrix.hpp:123-130).
 *   - lengths in nm, q in 1/nm, angles in radians unless a parameter says degrees.
 *
 * Error handling: every function returns HGX_OK (0) or a negative HGX_ERR_* code and never exits
 * the process; hgx_last_error(ctx) returns a message for the last failure on that context.
 * A context is not thread-safe; distinct contexts are independent (no global state).
 */
#ifndef HGX_H_
#define HGX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HGX_OK 0
#define HGX_ERR_INVALID_ARG (-1)
#define HGX_ERR_OOM (-2)
#define HGX_ERR_DEVICE (-3)
#define HGX_ERR_NO_QGRID (-4)
#define HGX_ERR_UNSUPPORTED (-5)
#define HGX_ERR_BAD_HANDLE (-6)

/* precision of a compute call */
#define HGX_PREC_FP64 0  /* plain: FP64 throughout */
#define HGX_PREC_FP32C 1 /* FP32 bulk, compensated accumulation, FP64 where cancellation needs it */

/* mesh record kinds */
#define HGX_MESH_APPROX8 0 /* 8 doubles per triangle: area, nx, ny, nz, cx, cy, cz, pad */
#define HGX_MESH_EXACT9 1  /* 9 doubles per triangle: v1, v2, v3 (counter-clockwise, outward) */

/* numeric form-factor methods */
#define HGX_FF_APPROX 0 /* centroid rule */
#define HGX_FF_EXACT 1  /* exact flat-triangle integral */

/* flags for hgx_ff_numeric */
#define HGX_FLAG_EXACT_REFERENCE_SIGN 1 /* exact method with the reference's exp(-iq.r) sign */
#define HGX_FLAG_ACC_KAHAN 2            /* FP32C: per-pair Kahan instead of FP32-block/FP64-carry */
#define HGX_FLAG_NO_MFMA 4              /* approx: per-pair VALU kernels even on a product grid */
#define HGX_FLAG_EXACT_CPU_THRESHOLD 8  /* exact method: the reference CPU's DBL_MIN degenerate-case
                                           threshold instead of its GPU path's 1e-13 */

/* flags for hgx_sf_lattice */
#define HGX_SF_COMPAT_LB2 1 /* reproduce the reference's lb[2]*lb[2] term (src/sf/sf.cpp:212) */
#define HGX_SF_FP32C 2      /* hgx_sf_lattice in FP32C: phases reduced in FP64, FP32 Dirichlet kernels (real q) */

/* analytic shapes */
#define HGX_SHAPE_SPHERE 0   /* dims: [radius] (+ [height] with HGX_SHAPE_TRUNCATED_SPHERE) */
#define HGX_SHAPE_CYLINDER 1 /* dims: [radius, height] */
#define HGX_SHAPE_BOX 2      /* dims: [xsize, ysize, height] */
#define HGX_SHAPE_CUBE 3     /* dims: [edge] */
#define HGX_SHAPE_PRISM3 4   /* dims: [edge, height]; uses tau, eta */
#define HGX_SHAPE_PRISM6 5   /* dims: [edge, height]; uses tau, eta */
#define HGX_SHAPE_PYRAMID 6  /* dims: [xsize, ysize, height, baseangle in degrees] */

#define HGX_MAX_NODES 40 /* distribution nodes per dimension */

typedef struct hgx_ctx hgx_ctx;

/* Distribution nodes of one analytic shape: dimension d has n[d] node values val[d][0..n-1] with
 * weights wt[d][0..n-1]; the form factor is sum over the node grid of prod_d wt * F(nodes).
 * (src/ff/ff_ana.cpp:235-284 builds such lists; the node grid per shape is the reference's loop
 * nest, e.g. src/ff/ff_ana_pyramid.cpp:133-143.) Host pointers. */
typedef struct hgx_shape_params {
  int n[4];
  const double* val[4];
  const double* wt[4];
  double tau; /* prism tilt angles (radians), src/ff/ff_ana_prism.cpp:94 */
  double eta;
  int flags;  /* HGX_SHAPE_COMPAT_* bits */
} hgx_shape_params;

/* Pyramid: evaluate the reference's k1..k4 (src/ff/ff_ana_pyramid.cpp:47-60), whose exponentials
 * have the opposite sign of the published truncated-pyramid form factor while the +-i factors keep
 * theirs; by default the published form (the volume integral of the pyramid) is used. */
#define HGX_SHAPE_COMPAT_PYRAMID 1
/* Sphere with heights (node dimension 1): the cap of height h cut from the top of the sphere and
 * resting on z = 0, as the reference's GPU path evaluates a sphere with a height parameter
 * (src/ff/gpu/ff_ana_sphere_gpu.cu:53-113): e^{i q_z (h - R)} times a 512-point trapezoid over
 * z in [R - h, R] of the disks 2 pi r^2 J1(r Re q_par) / (r Re q_par) e^{i q_z z}, r^2 = R^2 - z^2, each
 * disk 0 where |q_par| or r Re q_par is below 1e-13; every height at most the diameter of every
 * radius. Without the flag heights are not read, as the reference CPU path (src/ff/ff_ana_sphere.cpp:
 * 98-110). FP32C evaluates the cap in FP64 (its translation phase in FP32). */
#define HGX_SHAPE_TRUNCATED_SPHERE 2

/* ---- context ----------------------------------------------------------------------------- */

/* Creates a context on HIP device `device` with its own non-blocking stream.
 * Replaces the implicit global state of the reference's GPU classes (init_gpu,
 * include/init/gpu/init_gpu.cuh:25-32; __constant__ symbols src/ff/gpu/ff_ana_gpu.cu:30-32). */
int hgx_create(int device, hgx_ctx** out);
int hgx_destroy(hgx_ctx* ctx);
/* Use an external hipStream_t (passed as void*) for all further work; NULL is the null stream. */
int hgx_set_stream(hgx_ctx* ctx, void* hip_stream);
/* Go back to the context's own non-blocking stream. */
int hgx_use_own_stream(hgx_ctx* ctx);
int hgx_synchronize(hgx_ctx* ctx);
const char* hgx_last_error(const hgx_ctx* ctx);
const char* hgx_version(void);
/* Launch timing with HIP events recorded on the context's stream around every kernel launch made
 * by later calls: "main" kernels are the hot loops (triangle / node / lattice loops), "aux" the
 * finishing and setup kernels. hgx_profile_end synchronises the stream and returns the summed
 * durations (ms) and launch counts since hgx_profile_begin. Replaces the reference's per-call
 * CUDA-event timing (src/ff/gpu/ff_tri_gpu.cu:413-415, 508-512). */
int hgx_profile_begin(hgx_ctx* ctx);
int hgx_profile_end(hgx_ctx* ctx, double* main_ms, int64_t* main_launches, double* aux_ms, int64_t* aux_launches);

/* Device memory on the context's device, for callers without a HIP toolchain of their own (the C++
 * adapters, cgo/JNI-style bindings). The reference allocates and frees per call instead
 * (src/ff/gpu/ff_tri_gpu.cu:440-518). memcpy_d2h synchronises the context's stream; memcpy_h2d
 * waits for it first. */
int hgx_device_alloc(hgx_ctx* ctx, size_t bytes, void** ptr);
int hgx_device_free(hgx_ctx* ctx, void* ptr);
int hgx_memcpy_d2h(hgx_ctx* ctx, void* dst_host, const void* src_device, size_t bytes);
int hgx_memcpy_h2d(hgx_ctx* ctx, void* dst_device, const void* src_host, size_t bytes);

/* ---- q-grid (persistent per context) ----------------------------------------------------- */

/* QGrid::create, q-space region (src/model/qgrid.cpp:30-99), built on the device.
 * nbranch becomes 1 (qz real); call hgx_qz_extended for the DWBA branches. */
int hgx_qgrid_qspace(hgx_ctx* ctx, int ncol, int nrow, double qy_min, double qz_min, double qy_max,
                     double qz_max, double alpha_i, double k0);
/* Rectangular grid qx = 0, qy over columns, qz over rows (row 0 = highest qz). No reference
 * counterpart; the reference's region_qspace always derives qx from the Ewald sphere. */
int hgx_qgrid_explicit(hgx_ctx* ctx, int ncol, int nrow, double qy_min, double qz_min, double qy_max,
                       double qz_max, double alpha_i, double k0);
/* Flat detector: 2theta = atan((col - col_center) p / D), alpha_f = atan((nrow-1-row + 0.5 -
 * row_horizon) p / D), then the Ewald-sphere q of src/model/qgrid.cpp:76-84. */
int hgx_qgrid_detector(hgx_ctx* ctx, int ncol, int nrow, double pixel, double distance, double col_center,
                       double row_horizon, double alpha_i, double k0);
/* Explicit upload of a grid from host arrays: qx, qy [npix]; qz complex [nbranch*npix];
 * alpha_f [nrow] may be NULL (then hgx_fresnel is unavailable). npix = nrow * ncol. */
int hgx_qgrid_set(hgx_ctx* ctx, int ncol, int nrow, int nbranch, const double* qx, const double* qy,
                  const double* qz_ext, const double* alpha_f);
/* QGrid::create_qz_extended (src/model/qgrid.cpp:102-127): replaces qz by the 4 DWBA branches
 * [kzf-kzi, -kzf-kzi, kzf+kzi, -kzf+kzi] for a particle in a layer with 1-n^2 = dnl. */
int hgx_qz_extended(hgx_ctx* ctx, double k0, double alpha_i, double dnl_re, double dnl_im);
int hgx_qgrid_info(const hgx_ctx* ctx, int64_t* npix, int64_t* nq, int* nrow, int* ncol, int* nbranch);
/* *is_product = 1 when every q-point is exactly (qx[0], qy of its column, qz of its branch and row),
 * as for hgx_qgrid_explicit and its DWBA branches (checked on the device whenever the grid changes).
 * FP32C approximate-method form factors on such grids run as complex matrix products on the
 * matrix cores. No reference counterpart. */
int hgx_qgrid_is_product(const hgx_ctx* ctx, int* is_product);
/* Whether the current grid has the Ewald-sphere form of QGrid::create (src/model/qgrid.cpp:76-84) and
 * of detector grids: q = (a_i + c_i vx_j, c_i vy_j, qz(branch, row i)) for row values a_i, c_i and
 * column values v_j, fitted from the grid arrays to 1e-10 (1 + max|q|) whenever the grid changes.
 * dv (may be NULL) receives the largest half-range of c over a 64-row tile of the rows sorted by c,
 * times the largest |v_j| (1/nm): a mesh of radius R about its origin needs series orders for
 * x = dv R (see hgx_ff_numeric_path). No reference counterpart. */
int hgx_qgrid_is_ewald(const hgx_ctx* ctx, int* is_ewald, double* dv);
/* Copies the grid back to host arrays (any pointer may be NULL). */
int hgx_qgrid_get(hgx_ctx* ctx, double* qx, double* qy, double* qz_ext, double* alpha_f);

/* ---- meshes ------------------------------------------------------------------------------- */

/* Uploads a triangle mesh (host array of ntri records of the given kind) once; the reference
 * re-reads the mesh file on every call instead (src/ff/ff_num.cpp:972-1095). Exact9 meshes can
 * be evaluated with both methods (the approx records are derived as the OBJ reader does,
 * src/file/objectshape_reader.cpp:271-300); approx8 meshes only with HGX_FF_APPROX. */
int hgx_mesh_upload(hgx_ctx* ctx, int kind, int64_t ntri, const double* records, int* mesh_id);
int hgx_mesh_release(hgx_ctx* ctx, int mesh_id);
int hgx_mesh_info(const hgx_ctx* ctx, int mesh_id, int64_t* ntri, double* volume, double* radius);

/* ---- form factors ------------------------------------------------------------------------- */

/* Numeric (triangulated) form factor over the whole q-grid into ff[2*nq] (device).
 *  HGX_FF_APPROX replaces NumericFormFactorC/G::compute_approx_triangle
 *    (src/ff/cpu/ff_tri_cpu.cpp:200-246; src/ff/gpu/ff_tri_gpu.cu:395-521;
 *     signature include/ff/gpu/ff_num_gpu.cuh:55-58);
 *  HGX_FF_EXACT replaces compute_exact_triangle (src/ff/cpu/ff_tri_cpu.cpp:46-195;
 *    src/ff/gpu/ff_tri_gpu.cu:144-205). HGX_PREC_FP32C with HGX_FF_EXACT needs a real q-grid
 *    (HGX_ERR_UNSUPPORTED otherwise): each flat triangle is its centroid term times the second
 *    divided difference of exp(i x) at x_j = q.d_j (d_j: vertex minus centroid), in FP32 as the
 *    series 2 sum_n h_n(i x)/(n+2)! (128-triangle blocks with |q| max|d| <= 4) or from sinc-form
 *    first differences (beyond), plus an FP64 moment series at small |q|. It has no degenerate-case
 *    thresholds (HGX_FLAG_EXACT_CPU_THRESHOLD does not apply) and stays accurate on sliver
 *    triangles, where the closed form of the FP64 path (the reference's algorithm) cancels.
 * transvec (may be NULL) multiplies by exp(i q.t); the reference ignores the unit-cell location
 * for numeric shapes (src/ff/ff.cpp:44), so pass NULL to reproduce it.
 * At q = 0 the result is the mesh volume (the reference divides 0/0 there). */
int hgx_ff_numeric(hgx_ctx* ctx, int mesh_id, int method, int precision, int flags, const double rot[9],
                   const double transvec[3], double* ff);

/* kernels of hgx_ff_numeric (hgx_ff_numeric_path) */
#define HGX_PATH_PAIR 0    /* per-q-point VALU kernels (any grid; both methods) */
#define HGX_PATH_PRODUCT 1 /* approximate method on a product grid, on the matrix cores */
#define HGX_PATH_EWALD 2   /* FP32C approximate method on an Ewald-sphere grid (hgx_qgrid_is_ewald) that is
                              not a product grid, on the matrix cores with a series of *orders + 1 terms
                              in the per-row scale of q_par; used when x = dv R needs at most 3 orders */
/* Which kernels hgx_ff_numeric would run for these arguments on the current grid: *path = HGX_PATH_*,
 * *orders (may be NULL) the Ewald series orders (-1 on the other paths). No reference counterpart. */
int hgx_ff_numeric_path(const hgx_ctx* ctx, int mesh_id, int method, int precision, int flags, int* path,
                        int* orders);

/* hgx_ff_numeric without any device allocation: the same kernels and arguments, with their partial
 * sums and rotated records in a fixed per-device scratch array of 32 MiB instead of the context's
 * growing workspace (fewer triangle chunks where the array is short; FP32C results may differ from
 * hgx_ff_numeric's within the FP32C bar). No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops). Refuses with
 * HGX_ERR_INVALID_ARG a call whose rotated records alone exceed the array (FP32C on a product grid
 * with a rotation, meshes of more than about a million triangles). Every context on a device shares
 * that device's array, and it keeps the last call's contents: calls that overlap on one device (two
 * streams, two contexts, two host threads) race. No hipMalloc, no hipFree, no profiling events. */
int hgx_ff_numeric_noalloc(hgx_ctx* ctx, int mesh_id, int method, int precision, int flags, const double rot[9],
                           const double transvec[3], double* ff);

/* Record adjoint of the approximate numeric form factor. For L = Re sum_q W(q) F(q), F as computed by
 * hgx_ff_numeric(HGX_FF_APPROX) for the same mesh and rotation and W complex [2*nq] on the device,
 * writes grad[6 * ntri_uploaded] (device) = (dL/d(s n)_t, dL/dc_t) per triangle in the order the mesh
 * was uploaded (zero for triangles dropped as degenerate). Gradients of any mesh parameterisation
 * follow by the chain rule through (s n_t, c_t). Both precisions work on any grid; HGX_PREC_FP32C runs
 * the matrix cores on product grids and FP32 vector arithmetic elsewhere. No reference counterpart: the
 * reference differentiates by finite differences (src/analyzer/hipgisaxs_fit_lmvm.cpp:190-224). */
int hgx_ff_numeric_adjoint(hgx_ctx* ctx, int mesh_id, int precision, int flags, const double rot[9],
                           const double* W, double* grad);

/* Analytic form factor, replaces AnalyticFormFactorG::compute_<shape>
 * (include/ff/gpu/ff_ana_gpu.cuh:54-94) with the CPU formulas of src/ff/ff_ana_<shape>.cpp.
 * HGX_PREC_FP64: plain FP64. HGX_PREC_FP32C (real q only, else HGX_ERR_UNSUPPORTED): phases formed
 * and reduced to turns in FP64, FP32 sines and node sums, and, per q-point, the FP64 formula where
 * the prism and pyramid closed forms cancel (near qx^2 = 3 qy^2, 3 qy^2 = qx^2, qx qy = 0 and small
 * |q|); results in FP64. */
int hgx_ff_analytic(hgx_ctx* ctx, int shape, const hgx_shape_params* params, int precision,
                    const double rot[9], const double transvec[3], double* ff);

/* ---- structure factor --------------------------------------------------------------------- */

/* Finite-lattice structure factor into sf[2*nq], replaces StructureFactor::compute_structure_factor
 * and StructureFactorG::compute (src/sf/sf.cpp:85-284; include/sf/gpu/sf_gpu.cuh:30-70).
 * la, lb, lc: lattice vectors already multiplied by the grain scaling; repet: repetitions (values
 * < 1 become 1); center: grain centre; dw_sigma: Debye-Waller sigma (0 = off; not in the
 * reference). */
int hgx_sf_lattice(hgx_ctx* ctx, const double la[3], const double lb[3], const double lc[3],
                   const double repet[3], const double center[3], double dw_sigma, int flags,
                   const double rot[9], double* sf);

/* ---- Fresnel coefficients and DWBA -------------------------------------------------------- */

/* MultiLayer::propagation_coeffs (src/model/multilayer.cpp:110-155) on the grid's alpha_f rows:
 * fc[2*4*npix] = [Ti conj Tf, Ti conj Rf, Ri conj Tf, Ri conj Rf]. Layers below the vacuum are
 * given top to bottom (substrate last) by delta, beta, thickness; order is the particle's layer
 * (0 = vacuum). */
int hgx_fresnel(hgx_ctx* ctx, int nl, const double* delta, const double* beta, const double* thickness,
                int order, double k0, double alpha_i, double* fc);

/* I[p] (+)= |weight * sum_b fc_b sf_b ff_b|^2 over the grid's branches (inline in the reference:
 * src/sim/hipgisaxs_main.cpp:1084-1105). fc and/or sf may be NULL (treated as 1). */
int hgx_dwba_intensity(hgx_ctx* ctx, const double* fc, const double* sf, const double* ff, double weight,
                       int accumulate, double* I);

/* ---- batches of numeric form factors -------------------------------------------------------- */

/* One item of a batch: a mesh of the batch's context evaluated on the q-grid held by `grid` (any
 * context on the same device; NULL = the batch's context) under rotation rot, into ff (device,
 * [2 * nq of that grid]). */
typedef struct hgx_batch_item {
  hgx_ctx* grid;
  int mesh_id;
  double rot[9];
  double* ff;
} hgx_batch_item;

/* Prepares a batch: cuts it into work units (item, q-tile, run of triangle blocks) of similar cost,
 * orders them longest first and allocates the partial-sum buffer once. No reference counterpart: the
 * reference evaluates one shape per call (src/ff/gpu/ff_tri_gpu.cu:395-521). The contexts whose grids
 * the items name must outlive the batch; a grid replaced (any hgx_qgrid_* call) or extended
 * (hgx_qz_extended) after creation, or a mesh released, makes hgx_batch_run refuse with
 * HGX_ERR_INVALID_ARG before any launch: create the batch again. */
int hgx_batch_create(hgx_ctx* ctx, int nitems, const hgx_batch_item* items, int precision, int flags, int* batch);
/* Runs the batch: one launch over all work units plus one finishing launch. use_graph = 1 replays them
 * from a hipGraph captured on the first such call (no allocation, no host work per run). */
int hgx_batch_run(hgx_ctx* ctx, int batch, int use_graph);
int hgx_batch_info(hgx_ctx* ctx, int batch, int64_t* units, double* pairs);
int hgx_batch_release(hgx_ctx* ctx, int batch);
/* Work-unit shape of the context's later batches: q-points per thread (1, 2, 4, 8) and triangle
 * blocks (128 triangles each) per unit. Set by the autotuner. */
int hgx_tune_batch(hgx_ctx* ctx, int qpt, int blocks_per_unit);

/* ---- fused grain-ensemble intensity ------------------------------------------------------- */

/* A lattice of identical analytic particles seen by several grain orientations. */
typedef struct hgx_lattice_model {
  int shape;                  /* HGX_SHAPE_SPHERE (dims: radius) or HGX_SHAPE_CYLINDER (radius, height) */
  hgx_shape_params params;    /* distribution nodes and weights */
  double la[3], lb[3], lc[3]; /* lattice vectors already multiplied by the grain scaling */
  double repet[3];            /* repetitions (values < 1 become 1) */
  double center[3];           /* grain centre */
  double dw_sigma;            /* Debye-Waller sigma (0 = off) */
  double dn2_re, dn2_im;      /* contrast (1-n^2)_layer - (1-n^2)_particle (hipgisaxs_main.cpp:983) */
  int sf_flags;               /* HGX_SF_* */
  int ngrains;                /* >= 1 grain orientations */
  const double* rot;          /* host [ngrains][9], mq = rot_g q for both form and structure factor */
  const double* weight;       /* host [ngrains] grain weights w_g, or NULL for 1 */
} hgx_lattice_model;

/* I[p] (+)= sum_g |w_g dn2 sum_b fc_b F(R_g q_b) S(R_g q_b)|^2 over the grid's branches b: the grain
 * loop, form factor, structure factor and DWBA sum of HipGISAXS::run_gisaxs
 * (src/sim/hipgisaxs_main.cpp:925-1100) fused into one pass per pixel. fc (device, [2*nq]) may be
 * NULL for the Born approximation. FP32C requires real q. Runs as hgx_ensemble_intensity with one
 * element at the cell origin (so rotations about z share one form-factor evaluation). */
int hgx_lattice_intensity(hgx_ctx* ctx, const hgx_lattice_model* model, int precision, const double* fc,
                          int accumulate, double* I);

/* One analytic element of a unit cell: a shape with its distribution nodes, its rotation inside the
 * grain (the shape's xrot/yrot/zrot product, src/sim/hipgisaxs_main.cpp:986), its locations and its
 * contrast (1-n^2)_layer - (1-n^2)_shape. */
typedef struct hgx_ensemble_element {
  int shape;               /* HGX_SHAPE_* */
  hgx_shape_params params; /* host node arrays */
  double rot[9];           /* R_e: the element is evaluated at (R_g R_e) q */
  int nloc;                /* >= 1 */
  const double* loc;       /* host [nloc][3] locations t_l */
  double dn2_re, dn2_im;
} hgx_ensemble_element;

/* A lattice of unit cells of analytic elements seen by several grains, each with its own rotation,
 * centre and weight. */
typedef struct hgx_ensemble_model {
  int nelem;                       /* 1..64 */
  const hgx_ensemble_element* elem;
  double la[3], lb[3], lc[3];      /* lattice vectors already multiplied by the grain scaling */
  double repet[3];                 /* repetitions (values < 1 become 1) */
  double dw_sigma;                 /* Debye-Waller sigma (0 = off) */
  int sf_flags;                    /* HGX_SF_* */
  int ngrains;                     /* >= 1 */
  const double* rot;               /* host [ngrains][9]: R_g, for the structure factor */
  const double* center;            /* host [ngrains][3] grain centres, or NULL for the origin */
  const double* weight;            /* host [ngrains] weights, or NULL for 1 */
} hgx_ensemble_model;

/* I[p] (+)= sum_g |w_g sum_b fc_b S_g(R_g q_b) sum_e dn2_e F_e(R_g R_e q_b) sum_l exp(i (R_g R_e q_b).t_l)|^2:
 * the grain, element and location loops, structure factor and DWBA sum of run_gisaxs
 * (src/sim/hipgisaxs_main.cpp:925-1100) for any analytic unit cell in one pass per pixel. fc (device
 * [2*nq]) may be NULL (Born). HGX_PREC_FP64 equals hgx_ff_analytic + hgx_caxpy + hgx_sf_lattice +
 * hgx_dwba_intensity per grain to round-off; HGX_PREC_FP32C (real q, any cell) evaluates the form
 * factors as hgx_ff_analytic's FP32C and the Dirichlet kernels in FP32 on phases reduced in FP64 and
 * sums amplitudes in FP64. When every grain sees the cell alike
 * (spheres and cylinders whose grain rotations share their third row, e.g. rotations about z, at
 * locations every grain sees alike) the cell's form factor is evaluated once per q-point and branch
 * and only the structure factor per grain. Host tables are staged through pinned memory: no stream
 * synchronisation. */
int hgx_ensemble_intensity(hgx_ctx* ctx, const hgx_ensemble_model* model, int precision, const double* fc,
                           int accumulate, double* I);

/* Forward-mode derivatives of hgx_ensemble_intensity (FP64): I[p] as above (written, not accumulated)
 * and dI[j][p] = d/dt I[p](model + t tangents[j]) for ndir (1..64) directions, all in one launch.
 * tangents[j] has the model's counts (elements, shapes, node counts, locations, grains) and holds in
 * every real field the derivative of the model's field along direction j: node values and weights
 * (NULL arrays: zero), tau, eta, element rotations, locations, contrasts, lattice vectors, dw_sigma,
 * grain rotations, centres and weights (NULL: zero); repetitions and shapes are not differentiated.
 * fc is held fixed. The fit of HiG inputs uses it for the gradient of its objective in place of the
 * reference's central differences (src/analyzer/hipgisaxs_fit_lmvm.cpp:190-224). No counterpart in
 * the reference. dI is device [ndir][npix]. */
int hgx_ensemble_intensity_tangent(hgx_ctx* ctx, const hgx_ensemble_model* model, int ndir,
                                   const hgx_ensemble_model* tangents, const double* fc, double* I, double* dI);

/* ---- chi^2 and its gradient (forward / backward pair) ------------------------------------- */

/* One-grain model fitted by hgx_chi2_*: I(p) = s |sum_b fc_b dn2 F(m_b) S(m_b)|^2 with
 * m_b = R_x(tilt) rot0 q_b and theta = {radius, height, lattice constant a, tilt (rad), Debye-Waller
 * sigma, scale s}; the lattice vectors are a * (ua, ub, uc). */
typedef struct hgx_fit_model {
  int shape;                  /* HGX_SHAPE_CYLINDER (radius, height) or HGX_SHAPE_SPHERE (radius) */
  double ua[3], ub[3], uc[3]; /* lattice vectors for a = 1 (Lattice::construct_vectors) */
  double repet[3];
  double center[3];
  double dn2_re, dn2_im;      /* contrast */
  double rot0[9];             /* fixed grain orientation applied before the tilt */
  int sf_flags;               /* HGX_SF_* */
} hgx_fit_model;

/* Binds a model to device inputs (kept by pointer, caller-owned): fc [2*nq] or NULL (Born), the
 * observed image I_obs [npix], an optional 0/1 mask [npix]. The reference's objective is
 * HipGISAXSObjectiveFunction::operator() (src/analyzer/objective_func_hipgisaxs.cpp:245-332), whose
 * gradient TAO's LMVM takes by central differences (src/analyzer/hipgisaxs_fit_lmvm.cpp:190-224). */
int hgx_fit_plan_create(hgx_ctx* ctx, const hgx_fit_model* model, const double* fc, const double* I_obs,
                        const double* mask, int* plan);
int hgx_fit_plan_release(hgx_ctx* ctx, int plan);
/* Objective of the plan (HGX_METRIC_CHI2 unless hgx_fit_plan_metric chose another):
 * chi2 = sum_p mask_p (I_p(theta) - O_p)^2 / max(O_p, 1); I_model (device [npix]) may be NULL.
 * Synchronises the stream (the scalar is returned to the host). */
int hgx_chi2_fwd(hgx_ctx* ctx, int plan, const double theta[6], int precision, double* chi2, double* I_model);
/* The objective and its gradient (6 components, the height component is 0 for spheres) by the hand-written
 * adjoint: one pass over the pixels that recomputes the amplitudes instead of storing them (the
 * normalised metrics add one statistics pass before it). */
int hgx_chi2_bwd(hgx_ctx* ctx, int plan, const double theta[6], int precision, double* chi2, double grad[6]);

/* Objectives of a plan: the default chi^2, and the reference's distance functors
 * (include/analyzer/distance_functions.hpp, HiG fitting keyword "distancemetric") with O = I_obs, I the
 * model image and m the mask: */
#define HGX_METRIC_CHI2 0         /* sum m (I - O)^2 / max(O, 1) */
#define HGX_METRIC_L2 1           /* AbsoluteDifferenceSquare (:125-143): sum (m |O - I|)^2 */
#define HGX_METRIC_SQRT_UNIT_L2 2 /* "sqrt_unit_l2", SqrtUnitVectorNormL2DistanceSquare (:423-456):
                                     sum (m |sqrt O/||sqrt O|| - sqrt I/||sqrt I|| |)^2, ||x||^2 = sum m x^2 */
#define HGX_METRIC_SQRT_C_L2 3    /* "sqrt_c_l2", SqrtCNormL2DistanceSquare (:460-499):
                                     sum (m |sqrt O - c sqrt I|)^2, c = <sqrt I, sqrt O>_m / ||sqrt I||^2 */
#define HGX_METRIC_CBRT_UNIT_L2 4 /* "cbrt_unit_l2", CbrtUnitVectorNormL2DistanceSquare (:543-575): the
                                     unit-vector form over cube roots (hgx_image_distance only) */
#define HGX_METRIC_CBRT_C_L2 5    /* "cbrt_c_l2", CbrtCNormL2DistanceSquare (:579-616): the c-norm form
                                     over cube roots (hgx_image_distance only) */
int hgx_fit_plan_metric(hgx_ctx* ctx, int plan, int metric);
/* Objective values of n parameter vectors (host [n][6]) into out[n] (host): the population evaluation
 * of the reference's particle-swarm fit (src/analyzer/hipgisaxs_fit_pso.cpp:521-543), queued back to back
 * with one synchronisation. FP64. */
int hgx_chi2_fwd_many(hgx_ctx* ctx, int plan, int n, const double* thetas, int precision, double* out);
/* Bounded L-BFGS (memory 6) on the plan's objective with the adjoint gradient, in place of the
 * reference's TAO LMVM with central finite differences (src/analyzer/hipgisaxs_fit_lmvm.cpp:190-224):
 * theta (in: start, out: result) moves only in the components with free_params[j] != 0, kept within
 * [lo, hi] (steps measured in units of hi_j - lo_j, or of max(|theta_j|, 1) for an unbounded one); stops
 * after `iterations` or when max_j |projected dD/dtheta_j| (hi_j - lo_j) <= tol |D|. objective and
 * iterations_done may be NULL. Host loop, one value-and-gradient pass per trial point. */
int hgx_fit_lbfgs(hgx_ctx* ctx, int plan, double theta[6], const int free_params[6], const double lo[6],
                  const double hi[6], int iterations, double tol, double* objective, int* iterations_done);

/* ---- batched trapezoid-grating fit --------------------------------------------------------- */

/* Lines of trapezoidal cross-section in the y-z plane (base on z = 0) extruded along x over `length`,
 * repeated `periods` times along y with `pitch`. Each candidate's line is meshed on the device: the
 * cross-section perimeter in n_bottom + 2 n_side + n_top segments (a multiple of 16) and the length in
 * n_length segments (a multiple of 8), two triangles per quad, end caps fanned from the corner
 * centroid. Parameters theta = (top width, sidewall angle in rad, height, scale); the bottom width is
 * top + 2 height cot(angle). */
typedef struct hgx_grating_model {
  double length, pitch;
  int periods;
  int n_bottom, n_side, n_top, n_length;
  double dn2_re, dn2_im; /* contrast (1 - n^2) of the line material */
} hgx_grating_model;

/* I(theta) = scale |dn2 SF(qy) sum_b fc_b F_b(q_b; theta)|^2 with F the approximate-triangle form
 * factor of the line's mesh and SF = sin(N qy pitch / 2) / sin(qy pitch / 2); chi^2 = sum_p mask_p
 * (I_p - O_p)^2 / max(O_p, 1), leaving out pixels with |q_b| R < 0.1 on any branch (R: the line's
 * radius). The context's q-grid must be a product grid (FP32C on the matrix cores). fc [2*nq] or NULL
 * (Born), I_obs [npix] and mask [npix] (or NULL) are device arrays read by pointer. ncand candidates
 * are evaluated together; device memory is bounded by running them in groups. The plan keeps the
 * grid of its creation: after a grid is replaced (hgx_qgrid_*) or extended (hgx_qz_extended),
 * hgx_grating_eval, hgx_grating_lbfgs and hgx_grating_intensity refuse with HGX_ERR_INVALID_ARG
 * before any launch; create the plan again. No reference
 * counterpart: the reference fits with TAO/PSO and finite-difference gradients
 * (src/analyzer/hipgisaxs_fit_lmvm.cpp:190-224, src/analyzer/hipgisaxs_fit_pso.cpp). */
int hgx_grating_plan_create(hgx_ctx* ctx, const hgx_grating_model* model, int ncand, const double* fc,
                            const double* I_obs, const double* mask, int* plan);
int hgx_grating_plan_release(hgx_ctx* ctx, int plan);
/* Objective of the plan's candidates: HGX_METRIC_CHI2 (default), HGX_METRIC_L2, HGX_METRIC_SQRT_UNIT_L2 or
 * HGX_METRIC_SQRT_C_L2 (definitions at hgx_fit_plan_metric); the normalised forms add a statistics pass
 * and a per-candidate reduction on the device before the adjoint. The optimiser and hgx_grating_get
 * then work on that objective (reported where chi^2 is named). */
int hgx_grating_plan_metric(hgx_ctx* ctx, int plan, int metric);
/* Sets every candidate's theta (host [ncand][4]) and resets its optimiser state. */
int hgx_grating_set_theta(hgx_ctx* ctx, int plan, const double* theta);
/* chi^2 and its gradient (adjoint, one pass) for every candidate at its theta; use_graph replays the
 * evaluation from a hipGraph captured on first use. */
int hgx_grating_eval(hgx_ctx* ctx, int plan, int use_graph);
/* `iterations` L-BFGS iterations (memory 6, Armijo backtracking on the step) for every candidate, on the
 * device: one chi^2 + gradient evaluation per iteration, each iteration one hipGraph launch. */
int hgx_grating_lbfgs(hgx_ctx* ctx, int plan, int iterations);
/* Host copies of each candidate's accepted theta [ncand][4], chi^2 [ncand] and gradient [ncand][4]
 * (any may be NULL). Synchronises the stream. */
int hgx_grating_get(hgx_ctx* ctx, int plan, double* theta, double* chi2, double* grad);
/* Model image of candidate cand at its accepted theta into I [npix] (device). */
int hgx_grating_intensity(hgx_ctx* ctx, int plan, int cand, double* I);

/* ---- accumulation and image post-processing ------------------------------------------------ */

/* y[0..n) (+)= a x[0..n) over complex arrays (interleaved, device; accumulate = 0 overwrites y): the
 * element-and-location sum of the form factor, ff += dn2 * F (inline in the reference,
 * src/sim/hipgisaxs_main.cpp:1015). */
int hgx_caxpy(hgx_ctx* ctx, int64_t n, double a_re, double a_im, const double* x, int accumulate, double* y);
/* y[0..n) (+)= a x[0..n) over real arrays (device): the iratio-weighted sum of structure intensities
 * (src/sim/hipgisaxs_main.cpp:1351-1364). */
int hgx_axpy(hgx_ctx* ctx, int64_t n, double a, const double* x, int accumulate, double* y);
/* y[0..n) (+)= a x[0..n)^2 over real arrays (device; x may be y): the grain- and ensemble-correlated
 * structure sums GnE (sum_s iratio_s I_s^2) and GE ((sum_s iratio_s I_s)^2) of run_gisaxs
 * (src/sim/hipgisaxs_main.cpp:1231-1262, 1371-1398). */
int hgx_axpy_sq(hgx_ctx* ctx, int64_t n, double a, const double* x, int accumulate, double* y);

/* Smearing window of the reference: [i - h, i + h) with the kernel weight indexed from the window
 * start, which shifts the kernel at the image edges (include/numerics/convolutions.hpp:141-147). */
#define HGX_SMEAR_COMPAT 1
/* Separable Gaussian smearing of the image I [nrow][ncol] (device, in place), rows then columns, each
 * output the weight-normalised sum over the window of half-width h = ceil(6 sigma) pixels; replaces
 * Convolutions::convolution_gaussian_2d (include/numerics/convolutions.hpp:128-175), applied by
 * run_gisaxs when the smearing sigma exceeds DBL_MIN (src/sim/hipgisaxs_main.cpp:1409-1431). By
 * default the weight follows the distance |k - i| over [i - h, i + h]; HGX_SMEAR_COMPAT keeps the
 * reference's window. The weight table goes through pinned staging memory: no synchronisation. */
int hgx_gaussian_smear(hgx_ctx* ctx, int ncol, int nrow, double sigma, int flags, double* I);

/* Distance of a model image I to reference data O, both device [n], with an optional device mask
 * [n] of pixel weights (NULL: all 1), under one of the HGX_METRIC_* functors: the scalar that the
 * reference's DistanceMeasure objects compute in HipGISAXSObjectiveFunction::operator()
 * (src/analyzer/objective_func_hipgisaxs.cpp:284-287; include/analyzer/distance_functions.hpp). Two
 * passes over the images for the normalised forms (the norms, then the terms), one for chi^2 and L2;
 * block partials summed in a fixed order. Writes the host scalar *distance; synchronises the stream. */
int hgx_image_distance(hgx_ctx* ctx, int64_t n, int metric, const double* I, const double* O,
                       const double* mask, double* distance);
/* hgx_image_distance with the value left in device memory (*d_distance, one double) and no
 * synchronisation: distances of many images queued back to back and read back at once (the particle
 * swarm's generation, src/analyzer/hipgisaxs_fit_pso.cpp:521-543). No counterpart in the reference. */
int hgx_image_distance_async(hgx_ctx* ctx, int64_t n, int metric, const double* I, const double* O,
                             const double* mask, double* d_distance);
/* The distance (host *distance) and its derivatives along ndir tangent images dI (device [ndir][n]):
 * ddistance[j] = sum_p dD/dI_p dI[j][p] (host [ndir]), dD/dI from the functor's closed form (the
 * normalised forms through their norms and c); pixels whose term the functor drops, or with I <= 0
 * under the root forms, contribute 0. One synchronisation. No counterpart in the reference. */
int hgx_image_distance_tangent(hgx_ctx* ctx, int64_t n, int metric, const double* I, const double* O,
                               const double* mask, int ndir, const double* dI, double* distance, double* ddistance);

/* ---- host helpers ------------------------------------------------------------------------- */

/* AnalyticFormFactor::param_distribution (src/ff/ff_ana.cpp:235-284): node values and weights of a
 * shape parameter over [pmin, pmax] with nvalues points, weights for stat 0 (none: one weight 1),
 * 1 (uniform), 2 (gaussian pdf of mean, sd; mean replaced by the node midpoint when not finite).
 * Default (flags 0): exactly nvalues evenly spaced nodes, weights normalised to sum 1.
 * HGX_DIST_REFERENCE: the reference's float-accumulating do-while (ff_ana.cpp:249-254, which can
 * yield nvalues +- 1 nodes) and unnormalised weights (1 each, or the raw pdf, :271-272), as the
 * .hig driver uses to reproduce the reference's images. Writes at most cap nodes; returns the node
 * count or a negative error code. Host arithmetic only. */
#define HGX_DIST_REFERENCE 1
int hgx_param_distribution(double pmin, double pmax, int nvalues, int stat, double mean, double sd, int flags,
                           double* vals, double* wts, int cap);

/* Launch-shape knobs of the context's numeric form-factor kernels (q-points per thread for the FP64
 * and FP32C kernels, each 1, 2, 4 or 8, and the workgroup count the triangle split aims for). Set
 * by the autotuner. No reference counterpart (the reference hard-codes its tiles,
 * include/common/parameters.hpp:37-48). Every hgx_tune_* setting belongs to its context: contexts
 * do not share them, and batches and grating plans keep the setting of their creation. */
int hgx_tune_numeric(hgx_ctx* ctx, int qpt64, int qpt32, int target_wg);

/* Operand precision of the product-grid kernel. */
#define HGX_PRODUCT_FP32 0  /* FP32 operands on v_mfma_f32_32x32x2_f32 */
#define HGX_PRODUCT_F16X2 1 /* each FP32 operand split into two fp16 halves on v_mfma_f32_32x32x16_f16 */

/* Work-unit shape of the context's product-grid kernel (single calls, later batches and later
 * grating plans, whose forward pass uses the same kernel): triangle blocks (128 triangles each) per
 * unit, 3 (Gauss) or 4 real products per complex product (HGX_PRODUCT_F16X2 takes 4 only), and the
 * operand precision. Default (64, 4, HGX_PRODUCT_F16X2). Set by the autotuner. No reference
 * counterpart. */
int hgx_tune_product(hgx_ctx* ctx, int blocks_per_unit, int complex_mults, int operands);

/* Real products per complex product (4, default, or 3: Gauss's form) of the context's plain FP64
 * product-grid kernel (v_mfma_f64_16x16x4_f64; HGX_PREC_FP64 approximate form factors on product
 * grids, single calls and later batches). No reference counterpart. */
int hgx_tune_product64(hgx_ctx* ctx, int complex_mults);

/* Tile of the split-fp16 product-grid kernel. row_tiles: 32-row tiles per wave (1, 2 or 4); each
 * wave builds its column's operands once for all its rows. two_level = 1 adds each 128-triangle
 * block's sums to the work unit's sums after the block; 0 accumulates in place over the unit
 * (row_tiles 4 only: fewer registers, about 10x larger FP32 rounding error). shared_rows = 1 builds
 * the row operands once per workgroup of four waves (adjacent column tiles) and shares them through
 * LDS (row_tiles 2 or 4). Default (2, 1, 1). Set by the autotuner; applies to the
 * context's single calls and to batches and grating plans created afterwards. No reference
 * counterpart. */
int hgx_tune_product_tile(hgx_ctx* ctx, int row_tiles, int two_level, int shared_rows);

/* Schedule of the Ewald-grid kernel (HGX_PATH_EWALD): pipelined = 1 builds each step's column
 * operands during the previous step's matrix products (one wave per SIMD), 0 (default) in their own
 * step (two waves per SIMD). Read by each call of the context. No reference counterpart. */
int hgx_tune_ewald(hgx_ctx* ctx, int pipelined);

/* Operands of the FP32C product-grid record adjoint (hgx_ff_numeric_adjoint and the grating fit):
 * HGX_PRODUCT_F16X2 (default: split fp16 on v_mfma_f32_32x32x16_f16, 4 real products per complex
 * product) with col_tiles = 4 (default) or 2 32-column tiles per wave, or HGX_PRODUCT_FP32
 * (v_mfma_f32_32x32x2_f32; col_tiles ignored).
 * Read by each adjoint call of the context and at grating plan creation. No reference counterpart. */
int hgx_tune_adjoint(hgx_ctx* ctx, int operands, int col_tiles);

#ifdef __cplusplus
}
#endif

#endif /* HGX_H_ */
