/*
 * gsg.h -- C ABI of libgsg, the MI355X (gfx950) implementation of
 * GalerkinSparseGrids.jl's hot path: the D-dimensional DG derivative/flux operator
 * (multidim_derivative.jl) applied to the hierarchical ("vhier") coefficient vector
 * inside explicit time stepping (pdes.jl: wave, transport, Vlasov).
 *
 * Conventions
 *   - Every entry point returns int status: GSG_OK (0) or a negative GSG_E* code; the
 *     message of the last failure on the calling thread is gsg_last_error().
 *   - Vectors use exactly the vhier layout of the reference's V2Dref/D2Vref
 *     (dg_vmethods.jl:102-142): multi-levels in CartesianIndices order (dimension 1
 *     fastest) skipping cutoff(scheme, L, n) (schemes.jl:21-27), then cells, then modes,
 *     dimension 1 fastest.  Indices are 0-based here (Julia 1-based).
 *   - Host pointers are borrowed for the duration of the call.  Device pointers
 *     (the *_device entry points) must be HIP device allocations of the right size;
 *     `stream` is a hipStream_t, NULL meaning the null (default) stream, as the handle
 *     torch reports for its default stream.  A *_device call is stream-ordered: its
 *     work starts after everything already enqueued on `stream` and everything
 *     enqueued on `stream` afterwards starts after it; it returns without waiting.
 *   - A handle is single-caller: calls on one handle must not overlap.  The server
 *     (gsg_server_*) runs on library-owned threads and streams; its entry points are
 *     thread-safe.
 *   - Precision GSG_PREC_FP64 computes everything in double; GSG_PREC_FP32 applies the
 *     operator with float arithmetic and keeps the RK state and the stage accumulation
 *     in double: the stage input is float, the FP64 operator table is carried as two
 *     float tables (hi and the remainder lo) summed apart in float and added in double
 *     with the FP64 coefficients, and the wave system's gradient pass also reads u's
 *     float rounding remainder.
 */
#ifndef GSG_H
#define GSG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSG_OK 0
#define GSG_EINVAL -1    /* bad argument (shape, range, kind) */
#define GSG_ENOMEM -2    /* host or device allocation failed */
#define GSG_EDEVICE -3   /* HIP runtime error / no device */
#define GSG_ESTATE -4    /* wrong handle state (e.g. wait on unknown id) */
#define GSG_EFULL -5     /* admission refused: request exceeds the device budget */
#define GSG_ENONFINITE -6 /* a solve produced a non-finite value */

#define GSG_MAX_D 8

/* scheme: schemes.jl:21-27 */
#define GSG_SCHEME_SPARSE 0
#define GSG_SCHEME_FULL 1

/* operator kinds */
#define GSG_OP_DERIV 0      /* y = D_d x              : multidim_derivative.jl:61-65 D_matrix(D,d,k,n)   */
#define GSG_OP_GRAD_SUM 1   /* y = sum_d coef[d] D_d x : grad_matrix (:67-69) contracted with coef       */
#define GSG_OP_LAPLACIAN 2  /* y = sum_d D_d (D_d x)   : laplacian_matrix (:71-79)                      */
#define GSG_OP_WAVE 3       /* y = [[0,I],[L,0]] x, len 2N : wave_data RHS (pdes.jl:22-49)              */
#define GSG_OP_TRANSPORT 4  /* y = -sum_d a[d] D_d x   : transport RHS (pattern of pdes.jl:179,186,191)  */
#define GSG_OP_CSR 5        /* generic sparse matrix imported by gsg_csc_create                          */
#define GSG_OP_TRANSFORM 6  /* a sparse-grid basis transform (gsg_transform_create)                       */
#define GSG_OP_VLASOV 7     /* the Vlasov right-hand side of vlasov_evolve (gsg_vlasov_create)            */

/* bases of the Vlasov path (1d_nodal_basis.jl, multidim_nodal_basis.jl) */
#define GSG_BASIS_MODAL 0   /* the hierarchical (vhier) DG basis */
#define GSG_BASIS_NODAL 1   /* the hierarchical nodal basis, k in {2, 3, 5} */
#define GSG_BASIS_POINTS 2  /* values at the collocation points */

#define GSG_PREC_FP64 0
#define GSG_PREC_FP32 1

typedef struct gsg_op gsg_op;

typedef struct gsg_op_desc {
    int32_t D;         /* dimension, 1..GSG_MAX_D */
    int32_t k;         /* modes per cell (polynomial degree < k), 1..10 (K_max, 1d_dg_functions.jl:7) */
    int32_t n;         /* maximal level, 0..20 */
    int32_t scheme;    /* GSG_SCHEME_* */
    int32_t kind;      /* GSG_OP_* */
    int32_t dir;       /* GSG_OP_DERIV: direction, 1-based like the reference's d */
    int32_t precision; /* GSG_PREC_* */
    int32_t reserved;
    double coef[GSG_MAX_D]; /* GRAD_SUM: coefficients; TRANSPORT: advection velocity a */
} gsg_op_desc;

/* Last error message of the calling thread ("" if none). */
const char *gsg_last_error(void);

/* Library version string. */
const char *gsg_version(void);

/* Number of visible HIP devices (0 on a host without GPU); never fails. */
int gsg_device_count(void);

/* ---- host-side layout and operator tables (no GPU needed) ---------------------- */

/* dg_vmethods.jl:35-45 get_size(Val(D), k, n, scheme); < 0 on bad arguments. */
int64_t gsg_get_size(int32_t D, int32_t k, int32_t n, int32_t scheme);

/* dg_vmethods.jl:123-142 V2Dref: for each vhier index j, the 1-based (level, cell, mode)
 * tuples, written as levels[j*D+e], cells[j*D+e], modes[j*D+e]; len = get_size. */
int gsg_v2dref(int32_t D, int32_t k, int32_t n, int32_t scheme, int32_t *levels, int32_t *cells,
               int32_t *modes, int64_t len);

/* dg_basis.jl:185-191 dg_basis(k): k rows of 2k coefficients (x^i part then
 * sgn(x) x^i part), row = mode. */
int gsg_dg_basis(int32_t k, double *out);

/* 1d_derivative.jl:136-148 periodic_DLF_matrix(k, n; basis="hier") written dense,
 * row-major (k 2^n)^2.  Entries the 1-D coupling rule proves zero are written as 0. */
int gsg_dlf_1d(int32_t k, int32_t n, double *out, int64_t len);

/* multidim_derivative.jl:61-65 D_matrix(D, d, k, n; scheme) assembled as CSC with
 * 1-based colptr/rowval (the SparseMatrixCSC triplet).  Call once with nzval == NULL
 * to get *nnz, then again with arrays of that size (colptr has len+1 entries). */
int gsg_d_matrix_csc(int32_t D, int32_t d, int32_t k, int32_t n, int32_t scheme, int64_t *colptr,
                     int64_t *rowval, double *nzval, int64_t *nnz);

/* 1d_nodal_basis.jl:383-414 transform_1D(k, n, from, to) written dense, row-major
 * (k 2^n)^2, for (from, to) in {(MODAL, NODAL), (NODAL, MODAL), (NODAL, POINTS),
 * (POINTS, NODAL)}, entries below 1e-12 dropped as the reference's thresholds do;
 * k in {2, 3, 5}, n <= 12. */
int gsg_transform_1d(int32_t k, int32_t n, int32_t from, int32_t to, double *out, int64_t len);

/* x coordinates of the 1-D collocation points in point-index order (eval_points_1D,
 * 1d_nodal_basis.jl:180-211); len = k 2^n. */
int gsg_points_1d(int32_t k, int32_t n, double *x, int64_t len);

/* ---- kernel selection (autotuning) --------------------------------------------- */

/* No counterpart in the reference.  Two kernels apply the operator: the LDS fiber
 * kernel (grids with D <= 6, k <= 5, k^D <= 256 and D k^(D-1) <= 256) and the
 * scalar-plan gather kernel (every grid).  gsg_tune_set records which one runs for a
 * grid and precision; GSG_KERNEL_AUTO restores the built-in rule.  Takes effect for
 * operators, batches and servers created afterwards.  Thread-safe. */
#define GSG_KERNEL_AUTO 0
#define GSG_KERNEL_PLAN 1
#define GSG_KERNEL_TILE 2
/* FP32, segmented shapes (D <= 4, k <= 4, not D = k = 4): the segmented kernel with the
 * other number of coupled input cells moved per load round (4 where the rule takes 2,
 * 2 where it takes 4) in the RK stages; elsewhere it means GSG_KERNEL_TILE. */
#define GSG_KERNEL_TILE_ALT 3
int gsg_tune_set(int32_t D, int32_t k, int32_t n, int32_t precision, int32_t variant);
/* The variant that runs for (D, k, n, precision): GSG_KERNEL_PLAN, GSG_KERNEL_TILE or
 * GSG_KERNEL_TILE_ALT. */
int gsg_tune_get(int32_t D, int32_t k, int32_t n, int32_t precision);
/* Streams (1..8) a batch or server spreads its (D, k) classes over under the streams
 * schedule (0: the tuned count, a "lanes L" line of gsg_tune.txt, else 4).  Takes effect
 * for batches built afterwards. */
int gsg_tune_lanes(int32_t lanes);

/* ---- operators ---------------------------------------------------------------- */

/* Build an operator (host FP64 setup, then device upload).  Replaces the CSC that
 * D_matrix / grad_matrix / laplacian_matrix / wave_data return
 * (multidim_derivative.jl:57,68,78; pdes.jl:46). */
int gsg_op_create(const gsg_op_desc *desc, gsg_op **op);

/* Import an arbitrary sparse matrix given as the SparseMatrixCSC triplet with 1-based
 * colptr (n+1 entries) and rowval (nnz entries): the (m,n,colptr,rowval,nzval) format
 * vlasov_evolve writes (pdes.jl:149-155).  Applied by a CSR kernel. */
int gsg_csc_create(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                   const double *nzval, int64_t nnz, int32_t precision, gsg_op **op);

/* The D-dimensional sparse-grid basis transform `from` -> `to` (multidim_nodal_basis.jl
 * transform, :83-141; make_modal2point_matrices / make_point2modal_matrices :124-136)
 * as a device operator: y = T x.  The reference assembles T as a restricted Kronecker
 * product (a CSC matrix); here it is applied matrix-free, the 1-D transform along each
 * direction in turn, which is exact on the sparse grid because the transforms are
 * triangular in level.  FP64. */
int gsg_transform_create(int32_t D, int32_t k, int32_t n, int32_t scheme, int32_t from, int32_t to, gsg_op **op);

/* The same D-dimensional transform assembled as CSC with 1-based colptr/rowval (the
 * SparseMatrixCSC triplet that vlasov_evolve writes for m2n, n2p, p2n and n2m,
 * pdes.jl:145-163): exactly the matrix gsg_transform_create applies, i.e. the
 * reference's restricted Kronecker product without its 1e-12 drop of small products.
 * Call once with nzval == NULL to get *nnz, then again with arrays of that size (colptr
 * has len+1 entries).  Host only, no device needed. */
int gsg_transform_csc(int32_t D, int32_t k, int32_t n, int32_t scheme, int32_t from, int32_t to, int64_t *colptr,
                      int64_t *rowval, double *nzval, int64_t *nnz);

/* The right-hand side of vlasov_evolve's steprule (pdes.jl:131-227) on the phase space of
 * 2 Dx dimensions (x then v):  f' = n2m p2n ( -sum_d v_d . (n2p m2n D_d f)
 *                                              + sum_d F_d . (n2p m2n D_{Dx+d} f) )
 * with v_d the point values of get_xi_modal (basic_function_exact_coeffs.jl:12-18) and
 * F_point (Dx vectors of length N in the point basis, row-major [Dx][N]) or NULL for
 * free streaming.  GSG_PREC_FP32 applies the derivative in float; the transforms, the
 * pointwise products and the RK state are FP64.  gsg_op_apply evaluates the right-hand
 * side, gsg_op_rk4 integrates it with fixed-step RK4. */
typedef struct gsg_vlasov_desc {
    int32_t Dx;        /* spatial dimension; the phase space has 2 Dx <= GSG_MAX_D dimensions */
    int32_t k;         /* 2, 3 or 5 */
    int32_t n;         /* 0..12 */
    int32_t scheme;    /* GSG_SCHEME_* */
    int32_t precision; /* GSG_PREC_* */
    int32_t reserved;
} gsg_vlasov_desc;
int gsg_vlasov_create(const gsg_vlasov_desc *desc, const double *F_point, int64_t F_len, gsg_op **op);

/* Length of the vectors the operator maps (N, or 2N for GSG_OP_WAVE). */
int64_t gsg_op_len(const gsg_op *op);

/* y = op * x with host vectors of length len: the `RHS*x` / `Ds[d]*f` product of the
 * reference's time loop (pdes.jl:63,65,179-191).  Copies both ways per call. */
int gsg_op_apply(gsg_op *op, const double *x, double *y, int64_t len);

/* y = op * x with device vectors (double) of length len, stream-ordered. */
int gsg_op_apply_device(gsg_op *op, const double *x_dev, double *y_dev, int64_t len, void *stream);

/* Fixed-step classical RK4 of y' = op*y from t0 with step dt for nsteps steps; y_inout
 * is the host state (length len).  If save_every > 0 and snapshots != NULL, the state
 * after every save_every-th step is written to snapshots (row-major
 * [nsteps/save_every][len]).  State stays device-resident for the whole call.
 * Replaces the ode45/ode78 call on (t,x)->RHS*x (pdes.jl:62-68, 206-213). */
int gsg_op_rk4(gsg_op *op, double *y_inout, int64_t len, double t0, double dt, int64_t nsteps,
               int64_t save_every, double *snapshots);

/* Same as gsg_op_rk4 on a device-resident state (double, length len), stream-ordered on
 * `stream` (see the conventions above). */
int gsg_op_rk4_device(gsg_op *op, double *y_dev, int64_t len, double dt, int64_t nsteps, void *stream);

/* No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same computation as gsg_op_rk4_device, bit for bit, with the RK state kept in a
 * fixed device array of 64 MiB per device (held until the process ends) instead of the
 * handle's own buffers: the call allocates nothing and replays no graph (every step is
 * launched eagerly).  Operators of kind TRANSPORT, WAVE, GRAD_SUM or DERIV whose state
 * fits: 32 B (FP64) or 28 B (FP32) per state entry, plus D N 8 (FP64) or D N 4 (FP32)
 * bytes for the wave system's gradients; larger ones fail with GSG_EINVAL.  All calls on
 * one device share the array: calls that overlap in time on one device (on any streams,
 * from any threads or handles) race and must not be made. */
int gsg_op_rk4_device_noalloc(gsg_op *op, double *y_dev, int64_t len, double dt, int64_t nsteps, void *stream);

int gsg_op_destroy(gsg_op *op);

/* ---- batched independent solves (one launch per RK stage covers every solve) ---- */

typedef struct gsg_batch gsg_batch;

int gsg_batch_create(int32_t precision, gsg_batch **b);

/* Add one solve: operator kind must be GSG_OP_TRANSPORT or GSG_OP_WAVE; y0 is the host
 * initial state (len = N or 2N), advanced nsteps RK4 steps of size dt.  *slot is the
 * solve's index in the batch. */
int gsg_batch_add(gsg_batch *b, const gsg_op_desc *desc, const double *y0, int64_t len, double dt,
                  int64_t nsteps, int32_t *slot);

/* Upload, build tiles, run every solve to completion. */
int gsg_batch_run(gsg_batch *b);

/* Run `steps` RK4 steps of every unfinished solve (device-resident, no host copies). */
int gsg_batch_step(gsg_batch *b, int64_t steps);

/* Copy a solve's current state to the host. */
int gsg_batch_get(gsg_batch *b, int32_t slot, double *y, int64_t len);

/* Total DoFs (sum of state lengths) in the batch. */
int64_t gsg_batch_dofs(const gsg_batch *b);

/* Kernel timing (no counterpart in the reference; for benchmarks): while enabled, every
 * RK stage launch of the batch is bracketed by a HIP event pair on the batch's stream
 * and its duration is accumulated per (D, k, system, kernel variant) class. */
typedef struct gsg_kernel_stat {
    int32_t D, k, precision, sys; /* sys: 0 sum-form (transport), 1 wave */
    int32_t variant;              /* GSG_KERNEL_PLAN or GSG_KERNEL_TILE */
    int32_t reserved;
    int64_t launches;             /* kernel launches timed */
    double total_ms;              /* summed launch durations */
    double algo_bytes;            /* summed algorithmic HBM bytes of those launches */
    int64_t dofs;                 /* state DoFs of the class (per launch) */
} gsg_kernel_stat;

int gsg_batch_profile(gsg_batch *b, int32_t enable);

/* Launch schedule of an FP32 batch's RK stages (no counterpart in the reference):
 * GSG_SCHEDULE_STREAMS runs every (D, k) class's stage launches on one of four streams
 * so the classes overlap; GSG_SCHEDULE_FUSED runs one launch per stage (two for wave
 * systems) that covers every class served by the direct or segmented kernels, the other
 * classes launching after it; GSG_SCHEDULE_AUTO (the default) takes the tuned choice
 * (gsg_tune_schedule, or a "schedule S" line of gsg_tune.txt), else streams.  Both give
 * the same results bit for bit.  FP64 batches always use streams.  May be called at any
 * time; the next steps follow it. */
#define GSG_SCHEDULE_AUTO 0
#define GSG_SCHEDULE_STREAMS 1
#define GSG_SCHEDULE_FUSED 2
int gsg_batch_schedule(gsg_batch *b, int32_t schedule);
/* The process default for batches and servers left at GSG_SCHEDULE_AUTO. */
int gsg_tune_schedule(int32_t schedule);
int gsg_batch_kernel_stats(gsg_batch *b, gsg_kernel_stat *out, int32_t max, int32_t *count);

int gsg_batch_destroy(gsg_batch *b);

/* ---- serving: continuous batching of arriving solves ---------------------------- */

typedef struct gsg_server gsg_server;

typedef struct gsg_server_config {
    int64_t hbm_budget_bytes; /* admission budget for resident solves and the operator tables of
                                 their grids, each grid's tables counted once (0 = 75% of free HBM) */
    int32_t precision;        /* GSG_PREC_* */
    int32_t steps_per_tick;   /* RK4 steps each resident solve advances per scheduling tick */
    int32_t max_resident;     /* cap on resident solves (0 = unlimited) */
    int32_t reserved;
} gsg_server_config;

typedef struct gsg_server_stats {
    int64_t submitted, admitted, completed, queued, resident;
    int64_t ticks;
    double dof_updates;  /* sum of state_len * steps completed */
    double busy_seconds; /* device time spent in ticks */
    int64_t resident_bytes;
} gsg_server_stats;

#define GSG_REQ_QUEUED 0
#define GSG_REQ_RUNNING 1
#define GSG_REQ_DONE 2
#define GSG_REQ_FAILED 3

int gsg_server_start(const gsg_server_config *cfg, gsg_server **srv);

/* Submit a solve (copied; the caller may reuse y0 at once).  *id identifies it. */
int gsg_server_submit(gsg_server *srv, const gsg_op_desc *desc, const double *y0, int64_t len,
                      double dt, int64_t nsteps, int64_t *id);

int gsg_server_poll(gsg_server *srv, int64_t id, int32_t *status);

/* Block until request `id` is done; copy its final state to y (length len) and report
 * its latency (completion - submission) in nanoseconds.  Releases the request.  A failed
 * request returns its own code: GSG_ENONFINITE (non-finite state), GSG_EFULL (it cannot
 * fit even with nothing resident: earlier grids' operator tables hold the rest of the
 * budget), GSG_ESTATE otherwise (server stopped, failed tick). */
int gsg_server_wait(gsg_server *srv, int64_t id, double *y, int64_t len, int64_t *latency_ns);

int gsg_server_stats_get(gsg_server *srv, gsg_server_stats *st);

int gsg_server_stop(gsg_server *srv);

/* ---- projection of initial data on the device ----
 * vcoeffs_DG(D, k, n, f; scheme) (dg_vmethods.jl:149-179, through coeffs_DG and
 * hcubature, dg_methods.jl:85-106): the Galerkin coefficients <f, v_j> of every basis
 * function.  The reference integrates each coefficient with adaptive cubature
 * (rtol 1e-8); here every cell gets a tensor Gauss-Legendre rule of q nodes on each half
 * of each 1-D cell (P = 2q points per direction, P^D per cell), exact to rounding for
 * f polynomial of degree < 2q - k + 1 per half cell.  The caller evaluates f at the
 * nodes on the device, a range of cells at a time:
 *   gsg_proj_nodes    writes the node coordinates of cells [cell0, cell0 + ncell);
 *   gsg_proj_contract turns f at those nodes into the cells' k^D coefficients each,
 *                     written to out[cell0 k^D ...] (out of length get_size).
 * Cells are numbered in vector order (the coefficients of cell c are out[c k^D ...]).
 * All pointers are device pointers, stream-ordered on `stream` (NULL: the null stream).
 * gsg_proj_contract grows the handle's scratch (two buffers, up to ncell k (2q)^(D-1)
 * doubles) with a device synchronisation when a call needs more than before, so calls
 * are not meant for stream capture.  FP64. */
typedef struct gsg_proj gsg_proj;
int gsg_proj_create(int32_t D, int32_t k, int32_t n, int32_t scheme, int32_t q, gsg_proj **out);
int64_t gsg_proj_len(const gsg_proj *p);             /* get_size(D, k, n; scheme) */
int64_t gsg_proj_cells(const gsg_proj *p);           /* get_size / k^D */
int64_t gsg_proj_points_per_cell(const gsg_proj *p); /* (2q)^D */
/* X[d ld + i], d < D, i < ncell (2q)^D: point i = (cell - cell0) (2q)^D + j, node index j
 * direction 1 fastest. */
int gsg_proj_nodes(gsg_proj *p, int64_t cell0, int64_t ncell, double *X_dev, int64_t ld, void *stream);
/* The per-direction factors of the same nodes: X1[(d ncell + c) 2q + p] is the
 * direction-d coordinate of node p (0 <= p < 2q) of cell cell0 + c; len = D ncell 2q.
 * Node j of a cell has coordinates X1[.. p_d] with p_d the base-2q digits of j,
 * direction 1 fastest, so a caller can evaluate f by broadcasting instead of reading
 * gsg_proj_nodes' D (2q)^D coordinates per cell. */
int gsg_proj_nodes_1d(gsg_proj *p, int64_t cell0, int64_t ncell, double *X1_dev, int64_t len, void *stream);
/* fvals[i] = f at point i of gsg_proj_nodes (fvals_len = ncell (2q)^D). */
int gsg_proj_contract(gsg_proj *p, int64_t cell0, int64_t ncell, const double *fvals_dev, int64_t fvals_len,
                      double *out_dev, int64_t out_len, void *stream);
/* tensor_construct(D, k, n, vcoeff_array; scheme) (tensor_construct.jl:19-63) on the
 * device: out[j] = prod_d vecs[d][i_d(j)], with i_d(j) the 1-D vhier index of DoF j in
 * direction d, the product taken in direction order.  vecs: a host array of D device
 * pointers, each to the k 2^n coefficients of a 1-D projection (vcoeffs_DG(1, k, n, f_d)).
 * The coefficients of a separable f = prod_d f_d; exact, no quadrature. */
int gsg_proj_tensor(gsg_proj *p, const double *const *vecs_dev, double *out_dev, int64_t out_len, void *stream);
int gsg_proj_destroy(gsg_proj *p);

#ifdef __cplusplus
}
#endif

#endif /* GSG_H */
