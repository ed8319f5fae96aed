/*
 * saddlepoint_hip: an MI355X (gfx950) backend for the saddlepoint PDHG LP solver.
 *
 * C ABI. Each Backend method of the reference is one function here, one to
 * one, and keeps the method's argument meaning
 * (crates/saddlepoint-core/src/backend.rs:43-137). Handles are opaque; all
 * arrays are plain host pointers with explicit lengths. Every function returns
 * SP_OK (0) or an SP_ERR_* code and leaves a message for sp_last_error().
 *
 * Threading: a backend handle, and every problem and vector made from it, is
 * used by one host thread at a time (the reference's CudaBackend is not Sync
 * either, crates/saddlepoint-cuda/src/backend.rs:99). Work is issued on one
 * HIP stream in call order; functions that return scalars or host data block
 * until that data is ready, as in the reference.
 */
#ifndef SADDLEPOINT_HIP_H
#define SADDLEPOINT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SP_OK = 0,
    SP_ERR_UNAVAILABLE = 1, /* no device / code object (CudaError::Unavailable, backend.rs:37-39) */
    SP_ERR_HIP = 2,         /* a HIP runtime call failed (CudaError::Driver) */
    SP_ERR_SHAPE = 3,       /* length mismatch (the reference panics, sparse.rs:156-157) */
    SP_ERR_TOO_LARGE = 4,   /* beyond 32-bit sparse indices (CudaError::TooLarge, sparse.rs:60-62) */
    SP_ERR_INVALID = 5,     /* null handle or bad argument */
    SP_ERR_PROBLEM = 6,     /* ProblemError from validate() (problem.rs:147-194) */
    SP_ERR_NUMERICAL = 7,   /* SolveError::Numerical (solver.rs:230-232, 284-286) */
    SP_ERR_MEMORY = 8       /* out of device memory, or an admission budget exceeded */
};

typedef struct sp_backend sp_backend;
typedef struct sp_problem sp_problem;
typedef struct sp_vec sp_vec;

/* PrimalStats / DualStats / RayStats (core/src/backend.rs:14-41). */
typedef struct { double residual_sq, objective; } sp_primal_stats_t;
typedef struct { double residual_sq, objective; } sp_dual_stats_t;
typedef struct { double objective, violation_sq, norm_sq; } sp_ray_stats_t;

/* A CSR matrix on the host: CsrMatrix (core/src/problem.rs:16-24), usize
 * indices as int64. Columns sorted within each row. */
typedef struct {
    int64_t nrows, ncols, nnz;
    const int64_t *row_ptr; /* nrows + 1 */
    const int64_t *col_idx; /* nnz */
    const double *values;   /* nnz */
} sp_csr_view;

/* ScaledProblem (core/src/scaling.rs:37-50), the input of Backend::load. */
typedef struct {
    sp_csr_view a, at;
    const double *c, *lc, *uc, *lv, *uv; /* scaled: c, lv, uv length n; lc, uc length m */
    const double *row_scale;             /* m */
    const double *col_scale;             /* n */
    double bound_scale, obj_scale;
} sp_scaled_view;

/* LpProblem (core/src/problem.rs:131-145): the input of solve(). */
typedef struct {
    sp_csr_view a;
    const double *c, *lc, *uc, *lv, *uv;
    double c0;
    int maximize;
} sp_lp_view;

/* ---- lifecycle ----------------------------------------------------------- */

/* CudaBackend::new(ordinal) (cuda/src/backend.rs:129-140). Loads the gfx950
 * code object that sits next to this library. */
int sp_backend_create(int ordinal, sp_backend **out);
/* Free every solver and server created on `b` first: they use its stream and
 * kernels until they are freed. */
void sp_backend_destroy(sp_backend *b);
/* Backend::name (backend.rs:48). The string lives as long as the backend. */
const char *sp_backend_name(const sp_backend *b);
/* Message of the calling thread's last failed call on this backend ("" if
 * none, or if its last call succeeded). Messages are kept per thread, so a
 * call failing on one thread never changes the text another thread reads.
 * b may be NULL for failures of sp_backend_create. */
const char *sp_last_error(const sp_backend *b);

/* ---- problem -------------------------------------------------------------- */

/* Backend::load (backend.rs:51; cuda/src/backend.rs:229-248): uploads A', A'ᵀ,
 * the scaled vectors and the residual factors (scaling.rs:124-131). */
int sp_problem_load(sp_backend *b, const sp_scaled_view *p, sp_problem **out);
void sp_problem_free(sp_problem *p);

/* ---- vectors (backend.rs:53-55) --------------------------------------------- */

int sp_vec_upload(sp_backend *b, const double *data, int64_t len, sp_vec **out);
/* A zero vector, without a host copy (faster than uploading zeros). */
int sp_vec_zeros(sp_backend *b, int64_t len, sp_vec **out);
int sp_vec_download(sp_backend *b, const sp_vec *v, double *out, int64_t len);
int sp_vec_copy(sp_backend *b, const sp_vec *src, sp_vec *dst);
int64_t sp_vec_len(const sp_vec *v);
void sp_vec_free(sp_vec *v);

/* ---- compute: one function per Backend method (backend.rs:57-136) ----------- */

int sp_spmv(sp_backend *b, const sp_problem *p, const sp_vec *x, sp_vec *out);   /* :58 */
int sp_spmv_t(sp_backend *b, const sp_problem *p, const sp_vec *y, sp_vec *out); /* :60 */
int sp_primal_step(sp_backend *b, const sp_problem *p, const sp_vec *x, const sp_vec *aty, double tau,
                   sp_vec *x_t, sp_vec *x_bar);                                          /* :63-71 */
int sp_dual_step(sp_backend *b, const sp_problem *p, const sp_vec *y, const sp_vec *ax_bar, double sigma,
                 sp_vec *y_t);                                                           /* :74-81 */
int sp_halpern_step(sp_backend *b, sp_vec *z, const sp_vec *z_t, const sp_vec *anchor, double w,
                    double rho);                                                         /* :85-92 */
int sp_distance_sq(sp_backend *b, const sp_vec *a, const sp_vec *bv, double *out);      /* :95 */
int sp_sub(sp_backend *b, const sp_vec *a, const sp_vec *bv, sp_vec *out);              /* :97 */
int sp_norm_sq(sp_backend *b, const sp_vec *v, double *out);                            /* :99 */
int sp_dot(sp_backend *b, const sp_vec *a, const sp_vec *bv, double *out);              /* :101 */
int sp_scale(sp_backend *b, sp_vec *v, double alpha);                                   /* :103 */
int sp_primal_stats(sp_backend *b, const sp_problem *p, const sp_vec *x, const sp_vec *ax,
                    sp_primal_stats_t *out);                                               /* :106 */
int sp_dual_stats(sp_backend *b, const sp_problem *p, const sp_vec *y, const sp_vec *aty,
                  sp_dual_stats_t *out);                                                   /* :108 */
/* No counterpart in the reference. sp_dual_stats for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same kernels and result, bitwise, but
 * the per-workgroup partials and totals live in a static device array of
 * 8,200 doubles (64 KiB) that each GPU holds from the library's load until
 * the process ends, instead of the backend's buffers; the call allocates
 * nothing. Its reductions use at most 1,024 workgroups × 3 sums, so every
 * length fits; a launch shape needing more scratch is rejected with
 * SP_ERR_TOO_LARGE. All calls on one GPU share that array: calls that overlap
 * on one GPU (from several threads or streams) race. SP_ERR_UNAVAILABLE if
 * the array's address could not be taken on the backend's device. */
int sp_dual_stats_noalloc(sp_backend *b, const sp_problem *p, const sp_vec *y, const sp_vec *aty,
                          sp_dual_stats_t *out);
int sp_step_dual_residual_sq(sp_backend *b, const sp_problem *p, const sp_vec *x, const sp_vec *x_t,
                             const sp_vec *aty, const sp_vec *aty_t, double tau, double *out); /* :113-121 */
int sp_primal_ray_stats(sp_backend *b, const sp_problem *p, const sp_vec *dx, const sp_vec *adx,
                        sp_ray_stats_t *out);                                              /* :124-129 */
int sp_dual_ray_stats(sp_backend *b, const sp_problem *p, const sp_vec *dy, const sp_vec *atdy,
                      sp_ray_stats_t *out);                                                /* :131-136 */

/* ---- the PDHG driver over this backend ------------------------------------------ */

/* SolverParams (core/src/solver.rs:25-93) plus this backend's execution
 * options. Fill with sp_default_params. */
typedef struct {
    double tolerance;               /* 1e-6 */
    double infeasibility_tolerance; /* 1e-8 */
    uint64_t max_iterations;        /* 1e6 */
    double time_limit_s;            /* < 0: none */
    uint64_t check_every;           /* 64 */
    double step_size_factor;        /* 0.998 */
    double reflection;              /* 1.0 */
    double restart_sufficient, restart_necessary, restart_artificial; /* 0.2, 0.5, 0.36 */
    double pid_kp, pid_ki, pid_kd, pid_discount;                      /* 0.99, 0.01, 0, 0.3 */
    int64_t power_iterations;                                         /* 200 */
    int64_t geometric_mean_passes;                                    /* 12 (ScalingParams, scaling.rs:25-34) */
    int64_t ruiz_iterations;                                          /* 10 */
    int has_pock_chambolle;                                           /* 1 */
    double pock_chambolle_alpha;                                      /* 1.0 */
    int bound_objective_rescaling;                                    /* 1 */
    int termination;                                                  /* 0 projected, 1 step-implied */
    /* Execution options (no counterpart in the reference's SolverParams). */
    int fused; /* 1: between checks, one fused Aᵀ-SpMV+primal+Halpern and one fused
                  A-SpMV+dual+Halpern kernel per iteration (bitwise equal to 0) */
    int log_every_checks; /* 0 = silent (solver.rs:49-50) */
    int precision;        /* 0: FP64 throughout. 1: mixed (needs fused = 1): between checks the
                             matrix values, the SpMV sums and the iterates' displacements from
                             the FP64 iterate at the last check are FP32, while that base, the
                             Halpern anchors, every check (KKT, restart residuals, PID,
                             infeasibility) and the returned solution stay FP64 */
    int spmv_autotune;    /* 0: every matrix gets the default SpMV plan; 1: the autotuned rule
                             (sp_spmv_tuning_for) picks each matrix's plan from its row lengths */
    int gpu_scaling;      /* 1 (default): A is uploaded unscaled; the transpose and the scaling
                             passes run on the device (bitwise the host's result for
                             pock_chambolle_alpha = 1). 0: ScaledProblem::new on the host */
    int reorder_columns;  /* 1 (default): the driver relabels columns by decreasing nonzero count
                             before loading, so the most gathered entries share cache lines; each
                             row keeps its nonzero order and the solution comes back in the
                             original order. 0: the LP's own column order */
    int64_t mixed_stall_checks; /* precision = 1: checks without halving the relative KKT error after
                                   which the solve continues in FP64 (0, the default: no such switch) */
    int column_panels; /* A' rows with more than 64 nonzeros gather x̄ panel by panel, so that the
                          workgroups in flight share an L2-sized slice of it (sp_problem_set_panels):
                          k in [8, 30]: panels of 2^k columns; 0: off; -1 (default): k = 16 when A' has
                          at least 2^21 columns (a gathered vector of 16 MB or more), else off. With
                          panels the driver also relabels the rows, those long rows first (longest
                          first) and the others by their first column (y comes back in the LP's order;
                          the relabeling of reorder_rows, whatever its setting). Changes the
                          order of the long rows' sums and of sums over rows only. The serving loop
                          runs without panels */
    int hot_table; /* the A'ᵀ-side plain-iteration kernel keeps y's first H entries in LDS (H = 2048
                      in FP64, 4096 in mixed precision) and gathers only the others from memory, on
                      a persistent grid: -1 (default): when at least half of A'ᵀ's nonzeros gather
                      one of them (A' rows below H: skewed row lengths, rows relabeled) and A' has
                      at least 2^23 nonzeros; 1: the same without the size condition; 0: off.
                      Bitwise no change. The serving loop's batched kernels have no hot table */
    int reorder_rows; /* relabel the rows of A' as column panels do (rows of more than 64 nonzeros
                         first, longest first, then the others by their smallest column; y comes
                         back in the LP's order), so that each wave's gathers of y and x̄ fall on
                         fewer cache lines: -1 (default): for LPs of at least 2^20 nonzeros; 1:
                         always; 0: only with column panels (which need it). Changes the order of
                         sums over rows only. Served LPs are relabeled as their single solves */
} sp_params;

/* Status (solver.rs:95-102). */
enum { SP_OPTIMAL = 0, SP_PRIMAL_INFEASIBLE = 1, SP_DUAL_INFEASIBLE = 2, SP_ITERATION_LIMIT = 3, SP_TIME_LIMIT = 4 };

/* Solution (solver.rs:104-121) without x and y, which go to caller buffers. */
typedef struct {
    int status;
    double objective, dual_objective, primal_residual, dual_residual, gap;
    uint64_t iterations, restarts;
    double solve_time_s; /* from the start of run() (solver.rs:312, 436) */
    double total_time_s; /* scaling and upload included */
    double setup_time_s; /* scaling, transpose, upload */
    double final_omega;
    /* Mixed precision: plain iterations that ran in FP64 after the solver
       switched over (a near-certificate of infeasibility or a stall). */
    uint64_t mixed_fp64_iterations;
} sp_solution;

void sp_default_params(sp_params *p);

/* solve() (solver.rs:154-157) on this backend: validate, scale on the host,
 * load, run the restarted Halpern PDHG loop, unscale. x_out (n) and y_out (m)
 * may be NULL. */
int sp_solve(sp_backend *b, const sp_lp_view *lp, const sp_params *params, sp_solution *out, double *x_out,
             double *y_out);

/* The same driver as an object that runs in slices, for callers that time
 * or interleave iterations. sp_solver_create does solve()'s work up to the
 * loop (validate, scale, load, power iteration); sp_solver_advance runs at
 * most `iterations` more iterations and sets *done once the loop has ended;
 * sp_solver_result reports the Solution (status -1 while still running). */
typedef struct sp_solver sp_solver;
int sp_solver_create(sp_backend *b, const sp_lp_view *lp, const sp_params *params, sp_solver **out);
int sp_solver_advance(sp_solver *s, uint64_t iterations, int *done);
int sp_solver_result(sp_solver *s, sp_solution *out, double *x_out, double *y_out);
void sp_solver_free(sp_solver *s);
/* Measurement (needs params.fused = 1): average milliseconds of one launch of
 * the plain iteration's main kernel of one side (0: the A'ᵀ SpMV with the
 * primal step and Halpern update, 1: the A' SpMV with the dual step and
 * Halpern update; the FP32 forms in mixed precision), from two events around
 * `reps` back-to-back launches on copies of what the kernel writes. The
 * solver's iterates are unchanged. */
int sp_solver_kernel_time(sp_solver *s, int side, int reps, double *ms);
/* The solver's loaded problem (scaled, columns relabeled when the solver
 * relabels them), owned by the solver and valid until sp_solver_free: for
 * measurements such as sp_gather_replay_time on the matrices a solve uses. */
const sp_problem *sp_solver_problem(const sp_solver *s);

/* ---- measurement ------------------------------------------------------------- */

/* While enabled, every kernel launch on the backend's stream is bracketed by
 * two HIP events; enabling clears the totals. */
int sp_backend_profile(sp_backend *b, int enable);
/* Totals for kernel `index` (0-based, in load order): its name, summed
 * event time in milliseconds and launch count. SP_ERR_INVALID past the end. */
int sp_backend_profile_read(sp_backend *b, int index, char *name, int name_len, double *total_ms,
                            int64_t *launches);
/* Bandwidth probe (the counterpart of the reference's sp-bandwidth,
 * crates/saddlepoint-cuda/src/probe.rs:13-44, whose copy and triad kernels
 * come in f64 and f32, probe.rs:38-41): average milliseconds per launch of
 * kind 0 = copy of n doubles (16 bytes per element), 1 = triad c = a + s·b
 * over n doubles (24 bytes per element), 2 = n random 8-byte gathers from a
 * table of table_len doubles (indices streamed, results written: 12 bytes
 * per element plus the gather), 3 = copy of n floats (8 bytes per element),
 * 4 = triad over n floats (12 bytes per element). Kinds 5-8 have no
 * counterpart in the reference: how random gathers and a stream share the
 * memory system, in blocks of 2048 elements (n rounded up to whole blocks):
 * n 8-byte gathers from a table of table_len doubles (indices streamed) and
 * a stream of n 12-byte elements (a double and an int, like a matrix's
 * values and column indices); 5 = gathers only, 6 = stream only, 7 = gather
 * blocks and stream blocks in alternate workgroups, 8 = every workgroup a
 * stream block and then a gather block. A table_len past 2^31 - 1 (int32
 * indices) returns SP_ERR_TOO_LARGE. */
int sp_probe(sp_backend *b, int kind, int64_t n, int64_t table_len, int reps, double *ms_per_launch);
/* No counterpart in the reference: the ceiling a matrix's own access pattern
 * puts on its SpMV's gathers. Average milliseconds of one launch that loads,
 * for every nonzero of A' (transpose = 0) or A'ᵀ (1) of a loaded problem and
 * in its SpMV plan's order and work items (with column panels: the panel
 * matrix's, then the short rows'), the table entry its column index names
 * (a double, or a float with fp32 = 1) and nothing else. */
int sp_gather_replay_time(sp_backend *b, const sp_problem *p, int transpose, int fp32, int reps, double *ms);
/* Waits for all work queued on the backend's stream. */
int sp_backend_sync(sp_backend *b);

/* ---- reading LPs ------------------------------------------------------------------- */

/* An LP owned by the library, as read from a file. */
typedef struct sp_lp sp_lp;
/* read_mps_file (crates/saddlepoint-io/src/lib.rs:50-309): free-format MPS
 * (NAME, OBJSENSE, ROWS, COLUMNS with integer markers ignored, RHS, RANGES,
 * BOUNDS; first RHS/RANGES/BOUNDS set), assembled as from_triplets does
 * (problem.rs:33-77). SP_ERR_PROBLEM with "line N: ..." (sp_last_error(NULL))
 * on malformed input, SP_ERR_INVALID when the file cannot be read. */
int sp_lp_read_mps(const char *path, sp_lp **out);
/* A view of the LP's arrays (valid until sp_lp_free), ready for sp_solve,
 * sp_check or sp_server_submit. */
int sp_lp_view_of(const sp_lp *lp, sp_lp_view *out);
const char *sp_lp_name(const sp_lp *lp);
void sp_lp_free(sp_lp *lp);

/* CsrMatrix::transpose (problem.rs:88-109) on the device: a stable radix sort
 * of the nonzeros by column, the result the host's counting sort gives. The
 * input is an m×n CSR matrix; the caller allocates t_row_ptr[n+1],
 * t_col_idx[nnz], t_values[nnz]. (The device preconditioning and the checker
 * transpose this way.) */
int sp_transpose_gpu(sp_backend *b, int64_t nrows, int64_t ncols, const int64_t *row_ptr, const int64_t *col_idx,
                     const double *values, int64_t *t_row_ptr, int64_t *t_col_idx, double *t_values);

/* ---- the KKT checker ---------------------------------------------------------------- */

/* KktReport (crates/saddlepoint-check/src/lib.rs:10-25). */
typedef struct {
    double primal_objective, dual_objective;
    double constraint_violation, bound_violation, dual_violation;
    double relative_primal, relative_dual, relative_gap;
} sp_kkt_report;
/* check() (check/src/lib.rs:62-118) on the device: recomputes Ax and Aᵀy of the
 * original (unscaled) LP, which it uploads itself, and the violations, the
 * dual objective and the relative measures of a claimed (x[n], y[m]), with y
 * in the solver's sign convention. Shares nothing with a solve but the LP.
 * Sums are in a fixed tree order (not the reference's serial order). */
int sp_check(sp_backend *b, const sp_lp_view *lp, const double *x, const double *y, sp_kkt_report *out);

/* ---- SpMV plans and the autotuner ------------------------------------------------ */

/* No counterpart in the reference (cuSPARSE plans its own SpMV). How one
 * matrix's SpMV is cut into workgroups: stream workgroups take consecutive
 * rows with at most block_nnz nonzeros and max_rows rows; a row longer than
 * long_row gets workgroups of its own (4096 nonzeros each). The plan fixes
 * the summation order of rows that span several threads, so a matrix's plan
 * must be the same wherever its LP runs. */
typedef struct {
    int block_nnz; /* 1..2048 (default 2048) */
    int long_row;  /* 1..block_nnz (default 2048) */
    int max_rows;  /* 1..1024 (default 1024) */
} sp_spmv_tuning;
void sp_spmv_tuning_default(sp_spmv_tuning *out);
/* The plan the autotuned rule picks for a matrix with these row pointers (a
 * deterministic function of the row lengths; the rule table is generated by
 * saddlepoint_amd/autotune.py from timings of candidate plans). */
void sp_spmv_tuning_for(const int64_t *row_ptr, int64_t nrows, sp_spmv_tuning *out);
/* sp_problem_load with explicit plans for A' and A'ᵀ (NULL: the default). */
int sp_problem_load_tuned(sp_backend *b, const sp_scaled_view *p, const sp_spmv_tuning *a,
                          const sp_spmv_tuning *at, sp_problem **out);
/* Average milliseconds of one spmv (transpose = 0) or spmv_t (1) of a loaded
 * problem over `reps` launches, timed with HIP events on the backend's stream. */
int sp_spmv_time(sp_backend *b, const sp_problem *p, int transpose, int reps, double *ms);
/* No counterpart in the reference. Column panels for A' of a loaded problem:
 * the rows with more than 64 nonzeros are also kept as a panel matrix whose
 * rows are their parts in consecutive column ranges of 2^shift columns, stored
 * and run panel by panel (so the gathers of the workgroups in flight hit one
 * L2-sized slice of the input), and each row adds its panel sums in panel
 * order; the other rows run as before. Every A' product of the problem uses it
 * from then on (the same sums in every method). shift in [8, 30]; 0 removes
 * the panels. A matrix without such rows is left as it is. */
int sp_problem_set_panels(sp_backend *b, sp_problem *p, int shift);

/* ---- serving: continuous batching of many LPs on one GPU ---------------------- */

/* No counterpart in the reference (it solves one LP per call). LPs are
 * submitted at any time; host worker threads validate and scale them; a
 * chunk runs 64 iterations of every admitted LP (batched plain iterations,
 * per-LP checks) and admits and retires LPs at its edges, by HBM budget and
 * slot count. Each LP's result is bitwise the one it gets alone. Submitted
 * params need check_every = 64; max_iterations is rounded up to a multiple
 * of 64 (a limit above the largest multiple, such as UINT64_MAX for "no
 * limit", becomes that multiple). Not thread-safe per server except
 * sp_server_submit and sp_server_state, which may be called while another
 * thread runs chunks (each thread reads its own sp_last_error). */
typedef struct sp_server sp_server;
typedef struct {
    int64_t budget_bytes; /* device bytes the admitted LPs may hold (default 64 GiB) */
    int max_active;       /* most LPs admitted at once (default 256) */
    int prep_threads;     /* host threads validating and scaling (default 8) */
    int use_graphs;       /* replay the batched steps of a chunk from a hipGraph (default 1) */
    int64_t chunk_work;   /* which admitted LPs a chunk runs: the cheapest ones (cost nnz + m + n)
                             until their costs add up to chunk_work, plus any LP left out of the
                             last max_skipped chunks, so that a chunk holds enough work to fill the GPU and
                             small LPs are not paced by large ones. 0: the default (32,000,000);
                             negative: every admitted LP in every chunk. An LP's iterates do not
                             depend on which chunks it runs in. */
    int max_skipped;      /* an LP left out of this many chunks in a row runs in the next one
                             (0: the default, 8): the bound on how far chunk_work may pace a large
                             LP, which sets the latency of the largest LPs against the small ones' */
} sp_server_params;
/* Seconds on the server's clock (from its creation). */
typedef struct { double submitted, admitted, completed; } sp_job_times;
enum { SP_JOB_QUEUED = 0, SP_JOB_PREPARED = 1, SP_JOB_ACTIVE = 2, SP_JOB_DONE = 3, SP_JOB_FAILED = 4 };

void sp_default_server_params(sp_server_params *p);
int sp_server_create(sp_backend *b, const sp_server_params *params, sp_server **out);
void sp_server_free(sp_server *s);
/* Copies the LP; *id identifies it from then on. */
int sp_server_submit(sp_server *s, const sp_lp_view *lp, const sp_params *params, uint64_t *id);
/* One chunk; reports LPs admitted, waiting (submitted, not admitted) and finished so far. */
int sp_server_run_chunk(sp_server *s, int *active, int *waiting, int *finished);
int sp_server_state(sp_server *s, uint64_t id, int *state);
/* For a finished LP: the Solution, x (n) and y (m) (either may be NULL) and times.
 * A failed LP returns its error code. */
int sp_server_result(sp_server *s, uint64_t id, sp_solution *out, double *x, double *y, sp_job_times *times);
/* Drops a finished LP's stored result. */
int sp_server_forget(sp_server *s, uint64_t id);
double sp_server_clock(const sp_server *s);

/* ScaledProblem::new (scaling.rs:53-103) on the host, for callers that load
 * themselves. Arrays are caller-allocated: a_values[nnz] (the scaled A' in
 * the input's sparsity), c/lv/uv/col_scale[n], lc/uc/row_scale[m]. */
int sp_scale_problem(const sp_lp_view *lp, const sp_params *params, double *a_values, double *c, double *lc,
                     double *uc, double *lv, double *uv, double *row_scale, double *col_scale,
                     double *bound_scale, double *obj_scale);
/* The same on the device (the scaling passes of sp_params.gpu_scaling = 1),
 * read back into the same caller-allocated arrays; also at_values[nnz], the
 * scaled A'ᵀ in the transpose's sparsity. */
int sp_scale_problem_gpu(sp_backend *b, const sp_lp_view *lp, const sp_params *params, double *a_values,
                         double *at_values, double *c, double *lc, double *uc, double *lv, double *uv,
                         double *row_scale, double *col_scale, double *bound_scale, double *obj_scale);
/* ScaledProblem::unscale_primal (scaling.rs:114-116) on the device:
 * out = xs · col_scale / bound_scale entry by entry (the product, then a
 * correctly rounded division: bitwise the host's). All three vectors have the
 * LP's n entries. sp_solve and the solvers unscale this way from 2^20 entries
 * of x and y together. */
int sp_unscale_primal(sp_backend *b, const sp_vec *xs, const sp_vec *col_scale, double bound_scale, sp_vec *out);
/* ScaledProblem::unscale_dual (scaling.rs:119-121): out = ys · row_scale / obj_scale. */
int sp_unscale_dual(sp_backend *b, const sp_vec *ys, const sp_vec *row_scale, double obj_scale, sp_vec *out);

#ifdef __cplusplus
}
#endif
#endif
