/*
 * arbinterp.h -- C ABI of the MI355X-native Lekien-Marsden tricubic (3D) /
 * quadcubic (4D) interpolation library (libarbinterp.so, gfx950).
 *
 * The reference (DurhamDecLab/ARBInterp v1.8, src/ARBInterp/ARBInterp.py) is a
 * pure-Python class API with no FFI; each entry point below replaces one part
 * of that API, cited as ARBInterp.py:line.  The Python drop-in
 * (arbinterp_amd/ARBInterp.py) binds these through ctypes.
 *
 * Conventions
 *   - Every pointer argument named *_dev is HIP device memory; `stream` is a
 *     hipStream_t (NULL = the default stream).  Calls are stream-ordered; a
 *     call returns once its work is enqueued unless stated otherwise.
 *   - Return value: ARB_OK (0) or a negative ARB_E* code; arb_last_error()
 *     gives a thread-local message for the last failing call.
 *   - A handle is not thread-safe (neither is the reference object: Query
 *     mutates queryInds and the coefficient cache, ARBInterp.py:368-377).
 *     Its queries share the handle's device workspaces (sort buffers, packed
 *     output records, node tiles), so calls on one handle issued to
 *     different streams must not run at the same time: order the streams,
 *     or use one stream per handle.
 *   - All layouts are row-major.  Node/cell order is x fastest, then y, z (t):
 *     node = ix + nx*(iy + ny*(iz [+ nz*it]))  (ARBInterp.py:530-532,1266-1269);
 *     cell = cx + npx*(cy + npy*(cz [+ npz*ct])), npx = nx-3 (ARBInterp.py:368,1088).
 *   - Coefficient index i = ex + 4ey + 16ez [+ 64et] for u^ex v^ey w^ez [tau^et]
 *     (ARBInterp.py:111,730).
 */
#ifndef ARBINTERP_H
#define ARBINTERP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define ARB_OK      0
#define ARB_EINVAL (-1)   /* bad argument                                       */
#define ARB_ESHAPE (-2)   /* ncols is not d+1 or d+3: the reference sys.exit()s
                             (ARBInterp.py:103-104, :722-723)                     */
#define ARB_ENOMEM (-3)   /* device allocation failed                            */
#define ARB_EHIP   (-4)   /* a HIP runtime call failed                           */
#define ARB_EGRID  (-5)   /* rows do not form a complete grid with >= 4 nodes per
                             axis (the reference does not check, :528-568)       */
#define ARB_EINDEX (-6)   /* a query reduced to a flat cell index past the sentinel
                             column nc; the reference raises IndexError from
                             alphamask[queryInds] (ARBInterp.py:376,431,489)     */

/* ---- interpolation modes (ARBInterp.py:24-104, :643-723) --------------- */
#define ARB_MODE_SCALAR  0  /* N x (d+1) field: one interpolant, value + gradient;
                               sentinel column 0 (ARBInterp.py:24-32)           */
#define ARB_MODE_VECTOR  1  /* N x (d+3), mode='vector': 3 components; sentinel
                               column nc is 0, nc+1 is NaN (ARBInterp.py:35-48)  */
#define ARB_MODE_NORM    2  /* mode='norm': |B| value + gradient; sentinel NaN
                               (ARBInterp.py:49-58)                              */
#define ARB_MODE_BOTH    3  /* mode='both': components, |B|, grad |B|; sentinel
                               NaN (ARBInterp.py:59-74)                          */
#define ARB_MODE_VECTOR_NAN 4 /* no mode / unknown mode: components, sentinel
                               NaN (ARBInterp.py:75-102)                         */
#define ARB_MODE_F32_NORM 0x100 /* OR'ed into the mode of arb_field_create: the
                               rows hold float32 values, so |B| is formed in
                               float32 as numpy.linalg.norm does on a float32
                               array (ARBInterp.py:58,74,677,693)              */

/* ---- precision --------------------------------------------------------- */
#define ARB_PREC_FP64   0   /* plain: FP64 everywhere (separable coefficient
                               passes, table, evaluation)                       */
#define ARB_PREC_MIXED  1   /* FP64 cell reduction and coefficient passes; 3D
                               queries evaluate FP32 node-tile differences from
                               an FP64 per-block model; the table (4D, the
                               batch kernel, ARB_Q_TABLE) is formed in FP64 and
                               stored in FP32; FP32 evaluation                   */

/* ---- query flags ------------------------------------------------------- */
#define ARB_Q_NAN_ROWS  1   /* write NaN over every column of rows with a
                               coordinate outside [IntMin, IntMax] (as the
                               reference does: ARBInterp.py:350-355,1069-1076)      */
#define ARB_Q_CHECK     2   /* synchronise the stream and return ARB_EINDEX if any
                               row reduced past the sentinel                     */
#define ARB_Q_OUT_F64   4   /* mixed precision: write FP64 outputs (default FP32) */
#define ARB_Q_DIRECT    8   /* force the unsorted lane-per-query kernel          */
#define ARB_Q_SORTED   16   /* force the cell-bucketed path (LDS-staged blocks)  */
#define ARB_Q_SORTED_GLOBAL 32 /* bucketed order, coefficients read from HBM/L2  */
#define ARB_Q_TIMING   64   /* record per-stage HIP events; the call then waits
                               for its own work (see arb_last_timings)           */
#define ARB_Q_GROUP_ATOMIC 128 /* bucketed path: group queries with per-bucket
                               atomic ranks instead of the sort                  */
#define ARB_Q_PACKED  256   /* write each row's outputs as one aligned record of
                               arb_field_info.packed_width elements into
                               comps_dev: [comps(3)] [value, grad(d)], zero
                               padded; norms_dev/grads_dev are not used        */
#define ARB_Q_FUSED  1024   /* 3D bucketed path: form each block's
                               coefficients in the query kernel from node
                               values instead of staging them from the table.
                               No table is built or read; the results are
                               bitwise those of the table-reading paths      */
#define ARB_Q_SORT_CELL 512  /* bucketed path: order queries by bucket, then by
                               cell inside the bucket (default: bucket only)    */
#define ARB_Q_TABLE  2048   /* read each cell's coefficients from the table
                               even where the node tiles are the default
                               (mixed precision, 3D, direct and bucketed
                               paths: each query is evaluated from its 4x4x4
                               node values, no table is built or read; results
                               agree with the table path to FP32 rounding,
                               not bitwise)                                   */

#define ARB_Q_REDUCE_F32 4096 /* the field and the positions were float32 in the
                               caller (the values here are their exact FP64
                               widening): reduce cells in float32 and form the
                               power basis with float32 powers, as numpy does
                               for those dtypes (ARBInterp.py:358-373,
                               :493-503).  Uses the direct table kernel      */

/* ---- boundary extension (arb_field_set_boundary) ----------------------- */
#define ARB_BC_NONE      0  /* the reference's box: [2nd node, 2nd-last node]   */
#define ARB_BC_NEUMANN   1  /* ghost nodes by even reflection: zero normal slope
                               at the edge nodes                                 */
#define ARB_BC_DIRICHLET 2  /* ghost nodes by odd reflection about the edge node
                               (its value held, one-sided slope)                 */

typedef struct arb_field arb_field;

typedef struct arb_field_info {
    int32_t ndim;          /* 3 or 4                                            */
    int32_t mode;          /* ARB_MODE_*                                        */
    int32_t precision;     /* ARB_PREC_*                                        */
    int32_t ninterp;       /* interpolants in the table: 1, 3 or 4              */
    int32_t has_comps;     /* query writes comps (N x 3)                        */
    int32_t has_norm;      /* query writes norms (N) and grads (N x d)          */
    int32_t table_built;   /* 1 once arb_build_coeffs has run                   */
    int32_t coord_f32;     /* reserved, 0                                       */
    int64_t n[4];          /* nodes per axis (nPos+3)                           */
    int64_t npos[4];       /* cells per axis (nPos, ARBInterp.py:545)           */
    double  h[4];          /* |node0 - node1| per axis (ARBInterp.py:547-549)  */
    double  int_min[4];    /* 2nd node per axis (ARBInterp.py:551-553)          */
    double  int_max[4];    /* 2nd-last node per axis (ARBInterp.py:554-556)    */
    int64_t nc;            /* number of cells (ARBInterp.py:568)                */
    int64_t nnodes;
    int64_t table_bytes;   /* bytes of the coefficient table                    */
    int32_t packed_width;  /* elements per row with ARB_Q_PACKED (4 or 8)       */
    int32_t boundary;      /* ARB_BC_* (arb_field_set_boundary)                 */
    int32_t block[4];      /* cells per axis of one bucket of the bucketed
                              path (0 before its first use or arb_set_tuning) */
    int64_t tiles_bytes;   /* bytes of the node tiles (mixed 3D bucketed
                              path; 0 until its first query)                  */
    int32_t on_the_fly;    /* 1 when the coefficient table does not fit (device
                              memory or arb_set_table_limit): queries form
                              coefficients on the fly                         */
    int32_t reserved;
} arb_field_info;

/* Analyse a field and upload it (tricubic/quadcubic constructor minus the
 * coefficient fill: ARBInterp.py:10-104 -> getFieldParams :528-568,
 * makeAMatrix :107-175; 4D :629-723, :1264-1320, :726-878).
 * rows_dev: nrows x ncols FP64, row stride ld (>= ncols) elements, any row
 * order.  ndim 3|4; mode ARB_MODE_*; ncols must be ndim+1 (then mode is
 * forced to ARB_MODE_SCALAR, as the reference ignores it) or ndim+3.
 * Synchronises `stream` (the grid constants are needed on the host). */
int arb_field_create(const double* rows_dev, int64_t nrows, int32_t ncols, int64_t ld,
                     int32_t ndim, int32_t mode, int32_t precision,
                     void* stream, arb_field** out);

/* Grid constants (the reference's public attributes nPos, hx.., xIntMin..,
 * nc: ARBInterp.py:545-568). */
int arb_field_info_get(const arb_field* f, arb_field_info* out);

/* Extend the field by one ghost node layer per face so that the whole
 * sampled box [first node, last node] is interpolated (opt-in; no reference
 * counterpart: README.md:35 notes that queries at the edge return NaN and
 * plans Neumann or Dirichlet boundary conditions).  Every interpolated column
 * (components and |B| alike) is extended on its own, axis after axis (x, y,
 * z[, t]): Neumann v(-1) = v(1), v(n) = v(n-2); Dirichlet v(-1) = 2 v(0) -
 * v(1), v(n) = 2 v(n-1) - v(n-2).  Afterwards the field is the reference
 * algorithm on the extended grid: nPos = n - 1 cells per axis, IntMin /
 * IntMax = first / last node, h unchanged, cells numbered on the extended
 * grid.  Call it once, on a new field, before any table build or query;
 * arb_field_values, arb_field_set_values and arb_query_backward then return
 * ARB_EINVAL.  Synchronises `stream`. */
int arb_field_set_boundary(arb_field* f, int32_t bc, void* stream);

/* Fill the whole coefficient table (allCoeffs, ARBInterp.py:523-525,
 * :1260-1262, with calcCoefficients :570-603 / :1322-1355).  Idempotent.
 * When the table does not fit in device memory, or exceeds the limit of
 * arb_set_table_limit, it is not built: ARB_ENOMEM, and the field forms
 * coefficients on the fly from then on (arb_field_info.on_the_fly). */
int arb_build_coeffs(arb_field* f, void* stream);

/* Largest coefficient table the field may hold, in bytes (0: no limit other
 * than device memory; a held table above a new limit is released).  Fields
 * without a table answer every query on the fly, as the reference fills only
 * the cells that queries touch (ARBInterp.py:376-377, :489-490): 3D and 4D
 * table paths use fused bucketed kernels that form each bucket's
 * coefficients in LDS with the table build's arithmetic, so the results are
 * bitwise those of the table (mixed 3D fields need no table for their node
 * tiles); single points run as one-row batches.  arb_build_coeffs,
 * arb_table_copy, arb_table_load, arb_query_backward, arb_batch_create,
 * arb_push_rk4 and ARB_Q_REDUCE_F32 need the table and return ARB_ENOMEM. */
int arb_set_table_limit(arb_field* f, int64_t max_bytes);

/* Batch query (Query/rQuery, ARBInterp.py:177-211,344-521 / :880-914,1064-1258).
 * q_dev: n x ldq FP64 positions (first d columns used).  Outputs (device,
 * nullable when the mode has no such output): comps n x 3, norms n, grads
 * n x d; FP64 for ARB_PREC_FP64, FP32 for ARB_PREC_MIXED unless ARB_Q_OUT_F64.
 * cells_dev (nullable): int64 flat cell index per row, the reference's
 * queryInds (NaN -> nc).  Builds the table first if needed. */
int arb_query(arb_field* f, double* q_dev, int64_t n, int64_t ldq,
              void* comps_dev, void* norms_dev, void* grads_dev, int64_t* cells_dev,
              int32_t flags, void* stream);

/* ---- allocation-free query ------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Same arguments, flags and results as arb_query (the same
 * kernels: outputs are bitwise those of arb_query), but no call allocates:
 *   - the bucketed path's per-call workspace (sort keys and row
 *     indices, digit histograms, bucket bounds) is carved from one fixed
 *     device array of ARB_NOALLOC_SCRATCH_BYTES per GPU, held from the
 *     library's load until the process ends; separate output arrays are
 *     written directly (no packed-record workspace);
 *   - the field must already hold its coefficient table (table paths:
 *     arb_build_coeffs) or node tiles (mixed 3D: one arb_query with the same
 *     flags, n may be 0); they are rebuilt in place after
 *     arb_field_set_values.  Otherwise ARB_EINVAL;
 *   - n <= ARB_NOALLOC_MAX_ROWS, and the workspace of n rows over the field's
 *     buckets must fit the array (about 16 bytes per row, 32 on the mixed 3D
 *     node-tile path, plus 8 per bucket), else ARB_EINVAL with the sizes in
 *     the message;
 *   - ARB_Q_TIMING and ARB_Q_GROUP_ATOMIC are refused (ARB_EINVAL);
 *   - the field's device must be current.
 * Every call on one GPU shares that GPU's array and leaves its contents
 * behind (each call writes what it reads first), so calls that overlap on
 * one GPU -- on different streams, threads or handles -- race: order them. */
#define ARB_NOALLOC_SCRATCH_BYTES (80u << 20)
#define ARB_NOALLOC_MAX_ROWS (1 << 21)
int arb_query_noalloc(arb_field* f, double* q_dev, int64_t n, int64_t ldq,
                      void* comps_dev, void* norms_dev, void* grads_dev, int64_t* cells_dev,
                      int32_t flags, void* stream);

/* One point, host memory in and out: pt (d FP64); comps (3), norm (1), grad
 * (d), all FP64, for the outputs the mode has (else nullable); cell (int64,
 * nullable) = the flat cell index.  The point and the results pass through a
 * pinned host record owned by the field that the kernel reads and writes in
 * place: one launch and one synchronisation of `stream` per call, no device
 * allocation after the first call.  Returns ARB_EINDEX if the point reduces
 * to a cell past the sentinel (cell is written first).  Calls on one field
 * from several threads at once race on that record.  Replaces the per-point
 * chain of the reference's sQuery1/2/3 (ARBInterp.py:213-342, :916-1062). */
int arb_query_point(arb_field* f, const double* pt, double* comps, double* norm, double* grad,
                    int64_t* cell, void* stream);

/* Node values of the interpolated columns as stored, [ninterp][nnodes] FP64
 * in canonical node order: scalar f; vector Bx,By,Bz; norm |B|; both
 * Bx,By,Bz,|B| (the reference's Bn/Bx/By/Bz views of its sorted field,
 * ARBInterp.py:32,46-48,58,71-74). */
int arb_field_values(const arb_field* f, double* dst_dev, void* stream);

/* Replace the stored node values (same layout as arb_field_values) and mark
 * the coefficient table stale; the next query or arb_build_coeffs rebuilds
 * it.  Used by the differentiable path, where node values are trainable.
 * No reference counterpart (the reference field is fixed at construction). */
int arb_field_set_values(arb_field* f, const double* src_dev, void* stream);

/* Backward of the scalar interpolant's (value, gradient) outputs (modes
 * scalar, norm, both).  No reference counterpart.  For upstream gradients
 * gv_dev (n, FP64) and gg_dev (n x d, FP64), writes
 *   dq_dev   (n x d FP64, nullable): gv*grad U + Hess U . gg  (NaN rows where
 *            the position is outside the interpolation box);
 *   dval_dev (nnodes FP64, nullable): the gradient with respect to the stored
 *            scalar node values (last column of arb_field_values), computed
 *            as A^T applied to per-cell sums taken in query-index order and
 *            gathered per node in a fixed order (bitwise deterministic). */
int arb_query_backward(arb_field* f, const double* q_dev, int64_t n, int64_t ldq,
                       const double* gv_dev, const double* gg_dev, double* dq_dev,
                       double* dval_dev, void* stream);

/* ---- many fields, one launch ----------------------------------------------
 * No reference counterpart (the reference queries one field per call).  All
 * fields of a batch share ndim, mode and precision; their tables are built at
 * creation, and rebuilt by a batch query after arb_field_set_values.  The
 * batch keeps pointers to the fields: destroying one of them makes every
 * later arb_batch_query of the batch return ARB_EINVAL (destroy the batch
 * with arb_batch_destroy, in any order). */
typedef struct arb_batch arb_batch;

int arb_batch_create(arb_field* const* fields, int32_t nfields, void* stream, arb_batch** out);

/* Query every field of the batch in one persistent launch.  q_dev holds the
 * positions of field 0, then field 1, ... (counts[k] rows each, row stride
 * ldq); outputs use the same rows (comps n x 3, norms n, grads n x d, as in
 * arb_query).  Returns ARB_EINDEX (after synchronising) with ARB_Q_CHECK if a
 * row reduced past its field's sentinel. */
int arb_batch_query(arb_batch* b, double* q_dev, const int64_t* counts, int64_t ldq,
                    void* comps_dev, void* norms_dev, void* grads_dev, int32_t flags,
                    void* stream);

/* Work-item shape of the batched launch (autotuner hook): queries per work
 * item (multiple of 64), threads per workgroup (64..1024), workgroups per
 * compute unit.  0 keeps the current value. */
int arb_batch_set_tuning(arb_batch* b, int32_t chunk, int32_t threads, int32_t blocks_per_cu);

void arb_batch_destroy(arb_batch* b);

/* ---- particle push ------------------------------------------------------
 * Advance n particles (x_dev, v_dev: n x 3 FP64, updated in place) by nsteps
 * classical RK4 steps of dt with acceleration -grad U of the 3D field's
 * scalar interpolant (unit mass).  A particle whose stage position leaves
 * the interpolation box becomes NaN and stays NaN.  The steps are issued as
 * launches of steps_per_launch steps; with use_graph = 1 that launch
 * sequence is captured once into a hipGraph (re-captured when the arguments
 * change) and replayed on `stream`.  No reference counterpart (the
 * reference's trajectory code was removed in v1.8, CHANGELOG.md:8). */
int arb_push_rk4(arb_field* f, double* x_dev, double* v_dev, int64_t n, double dt,
                 int64_t nsteps, int32_t steps_per_launch, int32_t use_graph, void* stream);

/* Stage durations (ms, HIP events on the call's stream) of the last arb_query
 * made with ARB_Q_TIMING, five entries: [0] reduction + bucket keys (the
 * whole kernel on the direct path, reported in [3]), [1] grouping or radix sort
 * (bucket scan with ARB_Q_GROUP_ATOMIC), [2] scatter (ARB_Q_GROUP_ATOMIC), [3]
 * evaluation, [4] split of packed records into separate output arrays (0
 * when not used).  No reference counterpart (the reference has no timers);
 * used by the benchmark. */
int arb_last_timings(const arb_field* f, float* ms5);

/* A plain device-to-device copy of `bytes` (a multiple of 16; both pointers
 * 16-byte aligned) with one 16-byte load and store per thread: the
 * bandwidth a streaming kernel reaches on this GPU, measured by the
 * benchmark beside the evaluation kernel.  No reference counterpart. */
int arb_device_copy(void* dst_dev, const void* src_dev, int64_t bytes, void* stream);

/* Tuning of the bucketed query path (used by the autotuner): cells per block
 * along each axis (ndim entries, NULL = library default) and threads per
 * evaluation workgroup (64..256, 0 = default).  No reference counterpart. */
int arb_set_tuning(arb_field* f, const int32_t* block_cells, int32_t threads);

/* Copy cells [cell0, cell0+ncells) of the coefficient table as stored
 * ([cell][interp][K]; FP64 for ARB_PREC_FP64, FP32 for ARB_PREC_MIXED) to
 * dst_dev.  Cell nc is the sentinel column.  The reference keeps the same
 * numbers as alphax/alphay/alphaz/alphan[:, cell] (ARBInterp.py:42-74).
 * Builds the table first if needed. */
int arb_table_copy(arb_field* f, int64_t cell0, int64_t ncells, void* dst_dev, void* stream);

/* Upload cells [cell0, cell0+ncells) of the coefficient table from src_dev
 * (same layout and type as arb_table_copy).  A partial upload overlays the
 * built table (built first if needed).  From then on every query of the
 * field reads the table (the mixed 3D path no longer forms coefficients from
 * node values).  With arb_table_copy this saves and restores tables
 * (CHANGELOG.md:9 of the reference plans saving coefficients to a file). */
int arb_table_load(arb_field* f, int64_t cell0, int64_t ncells, const void* src_dev, void* stream);

/* Host-only: write the K x K operator matrix of makeAMatrix
 * (ARBInterp.py:107-175 / :726-878) into out (row-major, K = 4^ndim).
 * which: 0 = B (monomials and derivatives at the corners, = the reference's
 * examples/B_Matrix_{3D,4D}.csv), 1 = inv(B) (exact integers), 2 = D (central
 * differences in cuboid units, with the 4D row shift of ARBInterp.py:860),
 * 3 = A = inv(B).D.  Needs no GPU. */
int arb_operator_matrix(int32_t ndim, int32_t which, double* out);

/* Field ingestion (host only, no GPU): a numeric CSV file as
 * np.genfromtxt(path, delimiter=',') reads it, the way the reference's
 * examples load fields (examples/3D_ARBInterpExample.py:14).  '#' starts a
 * comment, blank lines are skipped, empty or non-numeric fields are NaN, and
 * every line must have the first line's column count (else ARB_ESHAPE).
 * arb_csv_shape counts rows and columns; arb_csv_read fills out
 * (nrows x ncols, row-major) on host threads. */
int arb_csv_shape(const char* path, int64_t* nrows, int32_t* ncols);
int arb_csv_read(const char* path, double* out, int64_t nrows, int32_t ncols);

/* Release the handle and its device memory. */
void arb_field_destroy(arb_field* f);

/* Thread-local message for the last failing call on this thread. */
const char* arb_last_error(void);

/* Library version string, e.g. "arbinterp-mi355x 0.1 (gfx950)". */
const char* arb_version(void);

#ifdef __cplusplus
}
#endif
#endif /* ARBINTERP_H */
