/* cfs_plume.h -- C ABI of the MI355X compressed-subspace step.
 *
 * Drop-in boundary for the reference's compressed linear algebra
 * (aadjones/CFS_PLUME, src/compression/COMPRESSION.h).  Every entry point is
 * plain C: opaque handles, raw pointers, sizes and int status codes (0 = ok,
 * < 0 = error; cfs_last_error() gives a message for the calling thread).
 * Device pointers are HIP device allocations; `stream` is a hipStream_t
 * (NULL = the default stream).  Calls on one basis handle must not overlap.
 */
#ifndef CFS_PLUME_H
#define CFS_PLUME_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFS_ABI_VERSION 1

/* status codes */
#define CFS_OK 0
#define CFS_ERR_ARG (-1)     /* bad argument (null pointer, range) */
#define CFS_ERR_IO (-2)      /* file could not be opened / read / written */
#define CFS_ERR_FORMAT (-3)  /* malformed component file */
#define CFS_ERR_HIP (-4)     /* HIP runtime error */
#define CFS_ERR_STREAM (-5)  /* the device decoder found an incomplete token stream */
#define CFS_ERR_NOMEM (-6)   /* host or device allocation failed */

/* device stream-error bit reported by cfs_check_stream_errors.  The loaders
 * parse every segment before upload and reject files whose segments do not
 * decode to exactly 512 values, have run lengths outside [2, 512] or end
 * inside a run (CFS_ERR_FORMAT), so this bit signals corrupted device data. */
#define CFS_STREAM_INCOMPLETE 0x2u

typedef struct cfs_basis cfs_basis;
typedef struct cfs_encoded cfs_encoded;

typedef struct {
  int32_t n_bits;        /* quantiser bits (COMPRESSION_DATA::get_nBits) */
  int32_t dims[3];       /* peeled grid X, Y, Z */
  int32_t padded[3];     /* padded to multiples of 8 */
  int32_t num_cols;      /* r */
  int32_t num_blocks;    /* nB per component */
  int32_t hazards;       /* boundaries where a singleton ends a segment and equals the next token */
  int64_t tokens[3];     /* int32 tokens per component file */
  int64_t device_bytes;  /* device memory held by the handle */
} cfs_basis_info;

/* Message for the last failing call on this thread ("" if none); with
 * cfs_abi_version, the library's own bookkeeping (no counterpart in the
 * reference, which prints and returns NULL or asserts). */
const char* cfs_last_error(void);
int cfs_abi_version(void);

/* ---- load -----------------------------------------------------------------
 * Replaces ReadBinaryFileToMemory x3 + MATRIX_COMPRESSION_DATA construction +
 * set_dampingArrayLists + dct_setup + init_cache
 * (COMPRESSION.cpp:1456-1594; MATRIX_COMPRESSION_DATA.h:59-122;
 *  SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:1488-1520, 1385-1393).
 * Reads the three component files (format v1), validates them, regroups the
 * tokens block-major and uploads them to the current HIP device.  Where the
 * reference returns NULL and prints, these return CFS_ERR_IO/FORMAT; where it
 * asserts on a token stream (COMPRESSION.cpp:1206-1207, 1267) or decodes a
 * short segment with trailing zeros, they return CFS_ERR_FORMAT. */
int cfs_basis_open_files(const char* path_x, const char* path_y, const char* path_z, cfs_basis** out);
/* Same from three in-memory file images (e.g. cfs_encoded_buffer). */
int cfs_basis_open_memory(const void* buf_x, size_t size_x, const void* buf_y, size_t size_y, const void* buf_z,
                          size_t size_z, cfs_basis** out);
int cfs_basis_close(cfs_basis* basis);
int cfs_basis_get_info(const cfs_basis* basis, cfs_basis_info* info);

/* ---- per-step operators (device pointers) ---------------------------------
 * v_full / field: full grid (X+2)(Y+2)(Z+2) cells, x fastest, interleaved
 * xyz, FP32 (VECTOR3_FIELD_3D layout, VECTOR3_FIELD_3D.cpp:69-85).
 * q: FP64[r]. */

/* PeeledCompressedProjectTransformNoSVD (COMPRESSION.cpp:2978-3016):
 * q = U^T v over the peeled cells, evaluated in the DCT domain. */
int cfs_project(cfs_basis* basis, const float* v_full, double* q, void* stream);

/* PeeledCompressedUnprojectTransform (COMPRESSION.cpp:2739-2900):
 * interior cells of field = U q; boundary cells are left untouched
 * (VECTOR3_FIELD_3D::setWithPeeled, VECTOR3_FIELD_3D.cpp:3415-3430). */
int cfs_unproject(cfs_basis* basis, const double* q, float* field, void* stream);

/* The per-step pair of stepPlume (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:477,490):
 * q_out = U_pre^T v_in and field = U_final q_in.  When pre == fin the two are
 * fused into one pass that decodes every segment once. */
int cfs_step(cfs_basis* pre, cfs_basis* fin, const float* v_in, double* q_out, const double* q_in, float* field,
             void* stream);

/* DecodeVectorField (COMPRESSION.cpp:2640-2670) for columns
 * [col0, col0+ncols): out holds ncols columns of 3*X*Y*Z FP32 values, peeled
 * and unpadded, interleaved xyz (the layout of one column of U). */
int cfs_decode_columns(cfs_basis* basis, int col0, int ncols, float* out, void* stream);

/* PeeledCompressedUnproject (COMPRESSION.cpp:2677-2711): the spatial variant
 * of cfs_unproject -- every column decoded spatially (DecodeVectorField) and
 * the interior of field set to sum_c q_c U_c (FP64 sums within a group of
 * columns, FP32 between groups); boundary cells untouched.  Columns are
 * decoded `group` at a time into `scratch` (device, group * 3 X Y Z floats;
 * null: allocated on the stream for the call).  q is device FP64[r]. */
int cfs_unproject_spatial(cfs_basis* basis, const double* q, float* field, float* scratch, int group, void* stream);

/* UntransformVectorFieldSVD (COMPRESSION.cpp:713-727) applied to ncols
 * decoded columns in place: columns holds ncols x 3*cells FP32 values in the
 * layout of cfs_decode_columns, V (device) holds ncols row-major 3 x 3 FP64
 * matrices; every cell's xyz triple u' of column c becomes V_c u'.  The
 * reference's live decode has the transform switched off (DecodeVectorField
 * :2667-2668, encoder :2312), so cfs_decode_columns does not call it. */
int cfs_untransform_svd(float* columns, int ncols, int64_t cells, const double* V, void* stream);

/* RunLengthDecodeBinary + ZigzagUnflatten (COMPRESSION.cpp:1198-1283,
 * 1427-1445) for blocks [block0, block0+nblocks) of (comp, col): out holds
 * nblocks x 512 int32 in u+8v+64w order.  Bit-exact check hook. */
int cfs_decode_block_ints(cfs_basis* basis, int comp, int col, int block0, int nblocks, int32_t* out, void* stream);

/* DecodeBlockWithCompressionData (COMPRESSION.cpp:1003-1057) for the same
 * range: FP32 coefficients int * (damp^gamma / s), u+8v+64w order. */
int cfs_decode_block_coeffs(cfs_basis* basis, int comp, int col, int block0, int nblocks, float* out, void* stream);

/* GetSubmatrixNoSVDSparse (COMPRESSION.cpp:1949-2158): the 3 x r rows
 * [start_row, start_row+3) of U, start_row a multiple of 3; out is FP64
 * column-major (out[comp + 3*col]) in device memory. */
int cfs_get_submatrix(cfs_basis* basis, int start_row, double* out, void* stream);

/* Synchronises, returns the CFS_STREAM_* bits raised since the last call and
 * clears them.  No counterpart in the reference, whose decoder asserts on a
 * malformed stream (COMPRESSION.cpp:1206-1207): a kernel records the fault
 * here instead. */
int cfs_check_stream_errors(cfs_basis* basis, uint32_t* bits);

/* ---- batched step: several bases, one persistent launch -------------------
 * No counterpart in the reference, which steps one basis pair per call
 * (stepPlume, SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:453-501): a batch steps
 * independent bases of different grids and column counts together.  Each
 * basis b gets q_out[b] = U_b^T v_in[b] and the interior of field[b] = U_b q_in[b]
 * (the fused cfs_step of every basis), in one launch whose workgroups are
 * split among the bases by work, plus one launch reducing every q.  The
 * bases must live on one device and outlive the batch; calls on one batch
 * must not overlap. */
typedef struct cfs_batch cfs_batch;
int cfs_batch_create(cfs_basis* const* bases, int n, cfs_batch** out);
/* Binds the device arrays of every basis (host arrays of n device pointers). */
int cfs_batch_bind(cfs_batch* batch, const float* const* v_in, double* const* q_out, const double* const* q_in,
                   float* const* field);
int cfs_batch_step(cfs_batch* batch, void* stream);
/* Autotuner: times every tile and launch shape (waves per workgroup 2 or 4
 * x workgroups per CU x work-split weight) `reps` times on the bound arrays
 * and keeps the fastest. */
int cfs_batch_tune(cfs_batch* batch, int reps, void* stream);
int cfs_batch_get_config(const cfs_batch* batch, int* grid, int* wg_per_cu, float* item_weight);
int cfs_batch_set_config(cfs_batch* batch, int wg_per_cu, float item_weight);
/* The tile shape: waves per workgroup (2 or 4; each wave takes one
 * (component, block) item at a time, the workgroup shares the damping
 * table, so two-wave workgroups halve the LDS per workgroup). */
int cfs_batch_get_shape(const cfs_batch* batch, int* waves_per_wg);
int cfs_batch_set_shape(cfs_batch* batch, int waves_per_wg);
/* Captures one whole batched step on `stream` (not the default stream) into
 * a HIP graph held by the batch; cfs_batch_graph_launch replays it.  Binding,
 * tuning or setting the configuration drops the graph. */
int cfs_batch_graph_capture(cfs_batch* batch, void* stream);
int cfs_batch_graph_launch(cfs_batch* batch, void* stream);
int cfs_batch_destroy(cfs_batch* batch);

/* ---- reduced advection and its learned closure ------------------------------
 * The reference advects in the subspace with cubature key cells
 * (reducedAdvectCompressionFriendly, SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:
 * 1781-1854, called from stepPlume :481): q' = sum_i A_i (sum_corners w U q).
 * cfs_reduced_advect evaluates the quantity that cubature approximates on
 * every cell: q_out = alpha * U^T A(U q_in) + forcing, with A the
 * first-order semi-Lagrangian self-advection VECTOR3_FIELD_3D::advect
 * (VECTOR3_FIELD_3D.cpp:2001-2077) over the full grid with backtrace dt0
 * (cells per unit velocity).  field_a / field_b: FP32 full-grid scratch
 * fields; forcing: FP64[r] or NULL.  It is the learned closure's fallback and
 * the reference its accuracy is measured against. */
int cfs_reduced_advect(cfs_basis* U, const double* q_in, double* q_out, float* field_a, float* field_b, float dt0,
                       double alpha, const double* forcing, void* stream);
/* cfs_reduced_advect without the two scratch fields.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The same
 * launches as cfs_reduced_advect with field_a / field_b taken from one
 * fixed device array that every GPU the library runs on holds until the
 * process ends (2 x CFS_NOALLOC_FIELD_FLOATS floats = 32 MiB); its address is
 * looked up on the first call on each device.  Calls nothing that allocates.
 * Full grids with 3 (X+2)(Y+2)(Z+2) > CFS_NOALLOC_FIELD_FLOATS are refused
 * with CFS_ERR_ARG.  Calls that overlap on one GPU race on the array. */
#define CFS_NOALLOC_FIELD_FLOATS 4194304
int cfs_reduced_advect_noalloc(cfs_basis* U, const double* q_in, double* q_out, float dt0, double alpha,
                               const double* forcing, void* stream);
/* ---- cubature reduced advection --------------------------------------------
 * The reference's per-step subspace advection, reducedAdvectCompressionFriendly
 * (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:1781-1854, accumAdvectRequests
 * :1859-1993): at K key cells x_i with weights w_i, backtrace by the velocity
 * U_pre(x_i) q, interpolate U_pre q trilinearly there and sum
 * q' = sum_i w_i U_after(x_i)^T a_i.  A cubature holds the key cells, weights
 * and the after-matrices w_i U_after(x_i)^T (FP64), built on the device from
 * the compressed basis `after` (readAdvectionCubature :1188-1327, which builds
 * them from U.prediffuse). */
typedef struct cfs_cubature cfs_cubature;
/* cells: peeled cell indices x + X (y + Y z) of `after`'s grid. */
int cfs_cubature_create(int K, const int32_t* cells, const double* weights, cfs_basis* after, cfs_cubature** out);
/* The `cubature` file: int32 K, int32 cells[K], FP64 weights[K]
 * (CUBATURE_GENERATOR_EIGEN::writeCubature, src/cubature/CUBATURE_GENERATOR_EIGEN.cpp:50-91;
 * read by readAdvectionCubature :1195-1232).  A missing file is CFS_ERR_IO,
 * a size that does not match K is CFS_ERR_FORMAT. */
/* The same with the after-basis dense (the reference reads U.prediffuse.matrix
 * and takes its rows at the key cells, cellBasisPeeled,
 * SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:1280-1292): rows (host) holds, for
 * key cell i, the 3 x r rows of U_after, entry (c, j) at i 3r + c + 3j;
 * X, Y, Z the peeled grid. */
int cfs_cubature_create_rows(int K, const int32_t* cells, const double* weights, int X, int Y, int Z, int r,
                             const double* rows, cfs_cubature** out);
int cfs_cubature_open_file(const char* path, cfs_basis* after, cfs_cubature** out);
int cfs_cubature_write_file(const char* path, int K, const int32_t* cells, const double* weights);
int cfs_cubature_get_info(const cfs_cubature* c, int* K, int* r);
int cfs_cubature_get_cells(const cfs_cubature* c, int32_t* cells, double* weights);
/* q_out = alpha * (cubature advection of q_in) + forcing (FP64[r] or NULL);
 * field: FP32 full-grid scratch (interior cells overwritten with U_pre q_in).
 * Three launches (unproject, key cells, reduction), no allocation. */
int cfs_cubature_advect(cfs_cubature* c, cfs_basis* pre, const double* q_in, double* q_out, float* field, float dt0,
                        double alpha, const double* forcing, void* stream);
/* The same step from the unprojected field U_pre q (FP32 full grid, interior
 * cells) instead of q: the key-cell interpolation and the sum of the after-
 * matrices only (accumAdvectRequests + reducedAdvectCompressionFriendly's
 * sum, :1781-1993), for callers that hold U_pre q already (a dense basis). */
int cfs_cubature_advect_field(cfs_cubature* c, const float* field, double* q_out, float dt0, double alpha,
                              const double* forcing, void* stream);
int cfs_cubature_destroy(cfs_cubature* c);

/* NNLS_SOLVER::solve (src/linearalgebra/NNLS.h:156-208, Lawson-Hanson) as the
 * cubature generator calls it (CUBATURE_GENERATOR_EIGEN.cpp:654-663):
 * x >= 0 minimising |A x - b|, A column-major m x n (host), at most
 * max_iter_factor * n inner iterations (the reference uses 5); *iterations
 * above that number means the cap was reached (x is then the last iterate,
 * still non-negative).  Host only. */
int cfs_nnls(const double* A, int m, int n, const double* b, int max_iter_factor, double* x, double* rnorm,
             int* iterations);

/* VECTOR3_FIELD_3D::advect of a full-grid field by itself (dims >= 4). */
int cfs_advect_field(const float* vel, float* out, int X, int Y, int Z, float dt0, void* stream);
/* VECTOR3_FIELD_3D::advect of a full-grid scalar field (VECTOR3_FIELD_3D.cpp:2082-2160,
 * the density and heat advection of advectHeatAndDensityStam,
 * SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:796-824): out = old sampled at the
 * backtrace of each cell along vel (FP32, X*Y*Z, x fastest); out != old. */
int cfs_advect_scalar(const float* vel, const float* old, float* out, int X, int Y, int Z, float dt0, void* stream);

/* Learned closure (no counterpart in the reference): an MLP r -> 256 -> 256 ->
 * 256 -> r with GELU, BF16 weights (FP32 master copies) and FP32
 * accumulation, predicting the scaled advection increment
 * (A(q) - q) / s_out from s_in * q.  Parameters are one FP32 array: for each
 * layer l = 0..3, W_l (K_l x M_l, row-major, input-major) then b_l (M_l),
 * K_0 = M_3 = r_padded (r rounded up to 64; padding rows/columns are zero),
 * other widths 256.  Network inputs/targets (forward, gradients, training)
 * are device FP32 arrays of n rows of r_padded (padding columns zero). */
typedef struct cfs_closure cfs_closure;
int cfs_closure_create(int r, int hidden, int max_batch, uint64_t seed, cfs_closure** out);
int cfs_closure_destroy(cfs_closure* c);
int cfs_closure_get_dims(const cfs_closure* c, int* r, int* r_padded, int* hidden, int64_t* n_params);
int cfs_closure_get_params(const cfs_closure* c, float* host_params);
/* Sets the parameters (host FP32) and resets the optimiser state. */
int cfs_closure_set_params(cfs_closure* c, const float* host_params);
int cfs_closure_set_scales(cfs_closure* c, const float* s_in, const float* s_out);
int cfs_closure_get_scales(const cfs_closure* c, float* s_in, float* s_out);
/* Optional FP32 linear skip L (host r x r, input-major: out_j += sum_k x_k L[k][j])
 * added to the network output by cfs_closure_step; NULL removes it.  The
 * network is then trained on the residual t - x L. */
int cfs_closure_set_linear(cfs_closure* c, const float* L);
/* Batched forward on MFMA (activations split BF16 hi + lo): y = MLP(x). */
int cfs_closure_forward(cfs_closure* c, const float* x, float* y, int n, void* stream);
/* Gradients (device FP32 array laid out like the parameters) of the
 * mean-squared error over the first r outputs, and the sum of squared errors
 * (device FP64), without an update. */
int cfs_closure_gradients(cfs_closure* c, const float* x, const float* t, int n, float* grads, double* loss,
                          void* stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on rows `rows[0..n)` of the
 * device training arrays x, t; the batch's sum of squared errors goes to
 * *loss_sum (device FP64, may be NULL). */
int cfs_closure_train_step(cfs_closure* c, const float* x, const float* t, const int32_t* rows, int n, float lr,
                           double* loss_sum, void* stream);
/* Closure step for n samples: q_out = alpha * (q_in + s_out * (MLP(x) + x L))
 * + forcing with x = s_in * q_in; q FP64 [n][r], forcing FP64 [n][r] or NULL. */
int cfs_closure_step(cfs_closure* c, const double* q_in, double* q_out, int n, double alpha, const double* forcing,
                     void* stream);

/* ---- the whole plume timestep on the device --------------------------------
 * stepPlume (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:453-501) without a host
 * round trip: buoyancy and vorticity confinement on the full grid
 * (FLUID_3D_MIC.cpp:491-593), q = U_pre^T v, the velocity border and the
 * density/heat advection (advectHeatAndDensityStam :796-824), the subspace
 * advection (cubature, learned closure or exact), q = D q then
 * q = P q + inverseProduct q (:848-854, :942-946), and the unprojection
 * v = U_final q.  Kernel launches only: a caller may capture a step (and
 * cfs_plume_add_smoke, addSmokeColumn FLUID_3D_MIC.cpp:655-703) in a HIP
 * graph.  Fields are the caller's full grids (velocity interleaved xyz,
 * density and heat scalar, FP32); q is FP64[r]. */
typedef struct cfs_plume_solver cfs_plume_solver;
typedef struct {
  double dt;             /* time step (_dt) */
  double dx;             /* cell size (_dx = 1 / max resolution) */
  double buoyancy;       /* _buoyancy */
  double vorticity_eps;  /* _vorticityEps after its resolution scaling; 0 turns confinement off */
  int open_x, open_y, open_z;  /* 1: open side, the border copies its neighbour
                                  (copyBorderX/Y/Z); 0: wall, the normal component is zeroed */
  double smoke_offset_x;       /* shift of the smoke column's axis, in cells (0: the reference's source;
                                  the learned closure's seeded runs move it) */
  double smoke_offset_z;
} cfs_plume_params;
#define CFS_PLUME_ADVECT_EXACT 0
#define CFS_PLUME_ADVECT_CUBATURE 1
#define CFS_PLUME_ADVECT_CLOSURE 2
/* The FLUID_3D_MIC constants for a full grid X x Y x Z and the plume
 * driver's walls (front/back/left/right walls, open top and bottom,
 * reducedCompressed3D.cpp:325-331). */
void cfs_plume_default_params(int X, int Y, int Z, cfs_plume_params* params);
/* The reduced coordinates may differ in size, as the reference's separate
 * bases do (U_preadvect r_pre columns, the advected coordinates r_adv --
 * U_prediffuse, the cubature's after-basis --, U_preproject r_proj, U_final
 * r_fin): damping is r_proj x r_adv (damping.peeled), ptof and
 * inverse_product are r_fin x r_proj (projected.ptof, inverseProduct), host
 * FP64 row-major; every dimension <= 256.  The exact advection and the
 * closure read U_preadvect coordinates; the exact advection writes them too
 * (r_adv = r_pre), a closure may write more (r_adv >= r_pre: its input is
 * q zero-padded to r_adv).  The q of
 * cfs_plume_step holds max(r_pre, r_fin) doubles; on return q[0, r_fin) are
 * the final coordinates. */
int cfs_plume_create(cfs_basis* pre, cfs_basis* fin, int r_adv, int r_proj, const double* damping, const double* ptof,
                     const double* inverse_product, const cfs_plume_params* params, cfs_plume_solver** out);
/* The uncompressed solver (SUBSPACE_FLUID_3D_EIGEN::stepPlume, the accuracy
 * harness's ground truth): the same step with dense bases, device FP64
 * 3 Npeel x r column-major (one mode per column, components interleaved),
 * owned by the caller; X, Y, Z the peeled grid.  Project and unproject are
 * dense products; the advection is the cubature (cfs_cubature_advect_field
 * on U_pre q), the closure or the exact advection. */
int cfs_plume_create_dense(const double* U_pre, int r_pre, const double* U_fin, int r_fin, int X, int Y, int Z,
                           int r_adv, int r_proj, const double* damping, const double* ptof,
                           const double* inverse_product, const cfs_plume_params* params, cfs_plume_solver** out);
/* Dense plumes only: the exact advection projects the advected field onto
 * this dense after-basis (device FP64, 3 Npeel x r_adv column-major, e.g.
 * U_prediffuse) instead of U_preadvect, q_adv = U_adv^T A(U_pre q) -- the
 * quantity the cubature approximates. */
int cfs_plume_set_exact_after(cfs_plume_solver* p, const double* U_adv);
/* Selects the subspace advection; the cubature or closure must outlive the plume. */
int cfs_plume_set_advection(cfs_plume_solver* p, int mode, cfs_cubature* cubature, cfs_closure* closure);
int cfs_plume_add_smoke(cfs_plume_solver* p, float* density, float* heat, void* stream);
int cfs_plume_step(cfs_plume_solver* p, float* velocity, float* density, float* heat, double* q, void* stream);
/* Replaces the step's parameters (e.g. a run's source offset and buoyancy). */
int cfs_plume_set_params(cfs_plume_solver* p, const cfs_plume_params* params);
/* Device pointers to the last step's subspace advection: its input (q after
 * the projection, r_pre values followed by zeros up to 256) and its output
 * (r_adv values) -- the pairs a learned closure is trained on. */
int cfs_plume_advection_state(const cfs_plume_solver* p, const double** q_pre, const double** q_adv);
int cfs_plume_destroy(cfs_plume_solver* p);

/* ---- the full-rank plume (snapshot source of the reduced bases) --------- */
/* FLUID_3D_MIC (src/integrators/FLUID_3D_MIC.cpp) on the device in FP64:
 * cfs_fluid_add_smoke = addSmokeColumn (:655-662), cfs_fluid_step = step
 * (:179-233: forces, advectStam :595-653, the peeled damping matrix
 * :1251-1307, project :314-402 with the MIC-preconditioned CG of
 * solvePoisson :918-1017), the snapshots of appendStreams (:1329-1480).
 * Fields live in the solver (cfs_fluid_fields: full grids X x Y x Z, x
 * fastest, velocity interleaved xyz, device FP64). */
typedef struct cfs_fluid cfs_fluid;
typedef struct {
  double dt;             /* _dt */
  double dx;             /* _dx = 1 / max resolution */
  double buoyancy;       /* _buoyancy */
  double vorticity_eps;  /* _vorticityEps (the drivers set it from the cfg "vorticity", default 0) */
  double damping_w;      /* w of buildPeeledDampingMatrixFull */
  double solver_eps;     /* _solverEps: CG stops when the residual maximum is at most this */
  int iterations;        /* _iterations: CG iteration cap */
  int boundaries[6];     /* front, back (z), left, right (x), top, bottom (y): 1 wall, 0 open
                            (FLUID_3D_MIC.cpp:100-118) */
} cfs_fluid_params;
/* Device pointers (any may be null) that cfs_fluid_step fills with the
 * peeled fields of appendStreams: velocity before advection, before
 * diffusion, before projection, final (3 (X-2)(Y-2)(Z-2) doubles each), the
 * pressure and the divergence ((X-2)(Y-2)(Z-2) each). */
typedef struct {
  double* preadvect;
  double* prediffuse;
  double* preproject;
  double* final_velocity;
  double* pressure;
  double* divergence;
} cfs_fluid_snapshots;
/* The FLUID_3D_MIC constructor's constants (FLUID_3D_MIC.cpp:29-56). */
void cfs_fluid_default_params(int X, int Y, int Z, cfs_fluid_params* params);
int cfs_fluid_create(int X, int Y, int Z, const cfs_fluid_params* params, cfs_fluid** out);
int cfs_fluid_fields(cfs_fluid* f, double** velocity, double** density, double** heat, double** pressure);
int cfs_fluid_add_smoke(cfs_fluid* f, void* stream);
/* One step; synchronises `stream` once per CG iteration (the CG control
 * flow runs on the host). */
int cfs_fluid_step(cfs_fluid* f, const cfs_fluid_snapshots* snapshots, void* stream);
/* CG iterations and residual maximum of the last pressure solve. */
int cfs_fluid_last_solve(const cfs_fluid* f, int* iterations, double* residual_max);
int cfs_fluid_destroy(cfs_fluid* f);

/* ---- moving box obstacle ----------------------------------------------
 * BOX (src/geometry/BOX.h, BOX.cpp:21-131): a box in the unit cube spinning
 * about the z axis with `period` and sliding its centre along x as
 * 0.25 sin(phi), phi = 2 pi t / translation_period.  A cell (full-grid index
 * x, y, z) is inside when its centre (x + 0.5, y + 0.5, z + 0.5) / (X, Y, Z)
 * (FLUID_3D_MIC::cellCenter :1704-1733), rotated back by theta about the
 * centre, lies within half_lengths; it then moves with omega x r + the slide
 * velocity (update_r, update_linearVelocity). */
typedef struct {
  double center[3];
  double original_center[3];
  double half_lengths[3];
  double period, translation_period;
  double theta, phi, time, dt;
} cfs_box;
/* BOX::BOX(center, lengths, period, translationPeriod) + set_dt (BOX.cpp:21-37). */
int cfs_box_init(const double center[3], const double lengths[3], double period, double translation_period,
                 double dt, cfs_box* box);
/* translate_center, spin, update_time (BOX.cpp:105-126, BOX.h:39-43): the
 * per-step advance of SUBSPACE_FLUID_3D_EIGEN.cpp:1715-1720. */
int cfs_box_advance(cfs_box* box);
/* FLUID_3D_MIC::setInitialVelocity (:744-763): the interior cells inside the
 * box take its velocity. */
int cfs_fluid_set_initial_velocity(cfs_fluid* f, const cfs_box* box, void* stream);
/* FLUID_3D_MIC::stepWithMovingObstacle (:236-310): the step without
 * buoyancy, with the box's cells stomped to its velocity
 * (setPeeledSparseMovingIOP :1645-1702) between the damping and the
 * projection.  iop_snapshot (device, 3 (X-2)(Y-2)(Z-2) doubles, may be null)
 * receives the peeled post-stomp velocity (appendStreams' IOP stream). */
int cfs_fluid_step_obstacle(cfs_fluid* f, const cfs_box* box, const cfs_fluid_snapshots* snapshots,
                            double* iop_snapshot, void* stream);
/* reducedSetMovingBox (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:779-790) with
 * the pair of buildObstacleMatrices (SUBSPACE_FLUID_3D_EIGEN.cpp:1699-1711):
 * q_out = R q + n, R = U_iop^T U_proj - U_iop^T C U_proj, n = U_iop^T nvec,
 * evaluated without R as iop_T_proj q + sum over the box's cells of
 * U_iop[c]^T (v_box(c) - U_proj[c] q).  U_iop, U_proj: device FP64 dense
 * bases, 3 X Y Z rows (peeled grid X x Y x Z), column-major; iop_T_proj:
 * device r_iop x r_proj row-major (U_iop^T U_proj); q (r_proj), q_out
 * (r_iop) device.  scratch: cfs_reduced_iop_scratch_bytes(X, Y, Z) bytes of
 * device memory, or null (allocated on the stream for the call). */
int64_t cfs_reduced_iop_scratch_bytes(int X, int Y, int Z);
int cfs_reduced_iop(const double* U_iop, int r_iop, const double* U_proj, int r_proj, const double* iop_T_proj,
                    int X, int Y, int Z, const cfs_box* box, const double* q, double* q_out, void* scratch,
                    void* stream);
/* The moving-obstacle pipeline of the reduced plume (initOutOfCoreIOP and
 * buildOutOfCoreMatricesIOP, SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:145-200,
 * SUBSPACE_FLUID_3D_EIGEN.cpp:1494-1659): U_iop and U_preproject (device,
 * 3 Npeel x r dense FP64 column-major, kept by the caller), U_iop^T
 * U_preproject (r_iop x r_proj) and the IOP basis' projection matrices
 * projected.ptof = U_final^T U_iop and inverseProduct (r_fin x r_iop each),
 * host row-major, copied. */
int cfs_plume_set_obstacle(cfs_plume_solver* p, int r_iop, const double* U_iop, const double* U_preproject,
                           const double* iop_T_preproject, const double* ptof_iop, const double* inverse_iop);
/* stepMovingObstacle (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:508-566): as
 * cfs_plume_step without buoyancy, with q = R q + n of the box at this step
 * (cfs_reduced_iop) between the diffusion and the pressure projection. */
int cfs_plume_step_obstacle(cfs_plume_solver* p, const cfs_box* box, float* velocity, float* density, float* heat,
                            double* q, void* stream);

/* ---- the basis builder's products (device FP64, column-major) ---------- */
/* C = A^T B (m x n, row-major, device): A rows x m (leading dimension lda),
 * B rows x n (ldb).  The snapshot Gram matrices of the SVD
 * (BIG_MATRIX::outOfCoreSVD, src/linearalgebra/BIG_MATRIX.cpp:276-420) and
 * the projections U^T (Op V) of buildOutOfCoreMatrices
 * (src/integrators/SUBSPACE_FLUID_3D_EIGEN.cpp:1355-1489, EIGEN::transposeProduct
 * src/linearalgebra/EIGEN.cpp:463-472); partial products are reduced in a
 * fixed order, so the result does not depend on scheduling. */
int cfs_gemm_tn(const double* A, int64_t rows, int m, int64_t lda, const double* B, int n, int64_t ldb, double* C,
                void* stream);
/* C = A M (rows x n, column-major, ldc): A rows x k (lda), M k x n row-major
 * on the device, k <= 512 (U = Q U_R of BIG_MATRIX.cpp:345-420). */
int cfs_gemm_nn(const double* A, int64_t rows, int k, int64_t lda, const double* M, int n, double* C, int64_t ldc,
                void* stream);
/* The sparse operators of the reduced projection, applied column by column
 * on the peeled grid of a full grid X x Y x Z (boundaries as in
 * cfs_fluid_params; dx, and w for the damping):
 *   CFS_OP_POISSON          buildFlatA (SUBSPACE_FLUID_3D_EIGEN.cpp:2166-2207), Np -> Np
 *   CFS_OP_VEL_TO_DIV       computeVelocityToDivergence (:1989-2100), 3 Np -> Np
 *   CFS_OP_PRESSURE_TO_VEL  computePressureToVelocity (:627-695), Np -> 3 Np
 *   CFS_OP_DAMPING          buildPeeledDampingMatrixFlat (:2107-2160), 3 Np -> 3 Np */
#define CFS_OP_POISSON 0
#define CFS_OP_VEL_TO_DIV 1
#define CFS_OP_PRESSURE_TO_VEL 2
#define CFS_OP_DAMPING 3
int cfs_apply_operator(int op, int X, int Y, int Z, const int* boundaries, double dx, double w, const double* in,
                       int64_t ld_in, int ncols, double* out, int64_t ld_out, void* stream);

/* ---- encoder (host) ---------------------------------------------------------
 * CompressAndWriteMatrixComponents (COMPRESSION.cpp:2271-2359) with the SVD
 * transform off, as on the live path (:2312).  U is FP64 column-major:
 * column c holds 3*X*Y*Z rows, row 3*(x + X*(y + Y*z)) + comp.  Each block:
 * edge-replicated padding, orthonormal DCT, scale by (2^(nBits-1)-1)/absMax,
 * gamma binary search (TuneGammaQuantized :841-916), rint quantisation,
 * zigzag, RLE.  `threads` <= 0 uses all hardware threads. */
int cfs_encode_matrix(const double* U, int X, int Y, int Z, int r, int n_bits, double percent, int max_iterations,
                      int threads, cfs_encoded** out);
/* One block of the same encoder: PreprocessBlock + TuneGammaQuantized +
 * EncodeBlock + ZigzagFlatten + RunLengthEncodeBinary (COMPRESSION.cpp:733-759,
 * 841-949, 1402-1421, 1112-1192) on F, a forward-DCT block of 512 FP64
 * coefficients in u+8v+64w order.  Writes s, gamma and *ntokens (<= 768)
 * tokens to `tokens`. */
int cfs_encode_block(const double* F, int n_bits, double percent, int max_iterations, double* s, double* gamma,
                     int32_t* tokens, int* ntokens);
/* The in-memory image of component file `comp` (0, 1, 2). */
int cfs_encoded_buffer(const cfs_encoded* enc, int comp, const void** data, size_t* size);
/* Writes <prefix>0, <prefix>1, <prefix>2 (the reference names U.<basis>.component0..2). */
int cfs_encoded_write(const cfs_encoded* enc, const char* prefix);
void cfs_encoded_free(cfs_encoded* enc);

#ifdef __cplusplus
}
#endif

#endif /* CFS_PLUME_H */
