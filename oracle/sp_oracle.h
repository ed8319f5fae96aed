/*
 * Oracle: a CPU restatement of saddlepoint's reference path, in plain C.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as the checker or
 * the timed CPU baseline. The product (saddlepoint_amd) never links it.
 *
 * What it restates (paths relative to the reference repository):
 *   - CsrMatrix::transpose / mul_vec / scale   crates/saddlepoint-core/src/problem.rs:88-126
 *   - ScaledProblem::new and helpers            crates/saddlepoint-core/src/scaling.rs:53-201
 *   - CpuBackend, all Backend methods           crates/saddlepoint-cpu/src/lib.rs:12-269
 *   - Solver::run, PID, restarts, KKT           crates/saddlepoint-core/src/solver.rs:154-480
 *   - check()                                   crates/saddlepoint-check/src/lib.rs:45-118
 *
 * Indices are int64 (the reference's usize). Reductions follow the CPU
 * backend's rule: a serial left fold below OR_PAR_THRESHOLD elements, and
 * above it a fixed-order combination of per-thread partial sums (rayon's
 * tree order is not reproducible; only the serial case is bitwise).
 */
#ifndef SP_ORACLE_H
#define SP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_PAR_THRESHOLD (32 * 1024) /* saddlepoint-cpu/src/lib.rs:12 */

typedef struct {
    int64_t nrows, ncols, nnz;
    int64_t *row_ptr; /* nrows + 1 */
    int64_t *col_idx; /* nnz */
    double *values;   /* nnz */
} or_csr;

/* minimize cᵀx + c0 s.t. lc ≤ Ax ≤ uc, lv ≤ x ≤ uv (problem.rs:131-145). */
typedef struct {
    or_csr a;
    const double *c, *lc, *uc, *lv, *uv;
    double c0;
    int maximize;
} or_lp;

typedef struct {
    int64_t geometric_mean_passes; /* 12 */
    int64_t ruiz_iterations;       /* 10 */
    int has_pock_chambolle;        /* 1 */
    double pock_chambolle_alpha;   /* 1.0 */
    int bound_objective_rescaling; /* 1 */
} or_scaling_params;

/* ScaledProblem (scaling.rs:37-50) plus the CpuProblem residual factors. */
typedef struct {
    or_csr a, at;
    double *c, *lc, *uc, *lv, *uv;
    double *row_scale, *col_scale;
    double bound_scale, obj_scale;
    double *primal_factor, *dual_factor; /* cpu/src/lib.rs:120-121 */
} or_scaled;

typedef struct {
    double tolerance;               /* 1e-6 */
    double infeasibility_tolerance; /* 1e-8 */
    uint64_t max_iterations;        /* 1e6 */
    double time_limit_s;            /* < 0: none */
    uint64_t check_every;           /* 64 */
    double step_size_factor;        /* 0.998 */
    double reflection;              /* 1.0 */
    double restart_sufficient, restart_necessary, restart_artificial;
    double pid_kp, pid_ki, pid_kd, pid_discount;
    int64_t power_iterations; /* 200 */
    or_scaling_params scaling;
    int termination; /* 0 projected, 1 step-implied */
    int log_every_checks; /* 0: silent; else the progress line of solver.rs:367-378 on stderr */
} or_solver_params;

enum { OR_OPTIMAL = 0, OR_PRIMAL_INFEASIBLE = 1, OR_DUAL_INFEASIBLE = 2, OR_ITERATION_LIMIT = 3, OR_TIME_LIMIT = 4 };
enum { OR_OK = 0, OR_ERR_PROBLEM = 1, OR_ERR_NUMERICAL = 2 };

typedef struct {
    int status;
    double objective, dual_objective, primal_residual, dual_residual, gap;
    uint64_t iterations, restarts;
    double solve_time_s; /* from the start of run(), as solver.rs:312 */
    double total_time_s; /* including scaling and load */
    double final_omega;
    double loop_time_s; /* the iteration loop alone (power iteration excluded) */
    char error[128];
} or_solution;

typedef struct {
    double primal_objective, dual_objective;
    double constraint_violation, bound_violation, dual_violation;
    double relative_primal, relative_dual, relative_gap;
} or_kkt_report;

typedef struct { double residual_sq, objective; } or_primal_stats;
typedef struct { double residual_sq, objective; } or_dual_stats;
typedef struct { double objective, violation_sq, norm_sq; } or_ray_stats;

/* threads <= 1 keeps every pass serial; otherwise passes at or above the
 * threshold use that many OpenMP threads. */
void or_set_threads(int threads);
int or_get_threads(void);

void or_default_scaling_params(or_scaling_params *p);
void or_default_solver_params(or_solver_params *p);

/* CSR helpers (problem.rs). `out` arrays are caller-allocated. */
void or_csr_transpose(const or_csr *a, int64_t *row_ptr_out, int64_t *col_idx_out, double *values_out);
void or_csr_mul_vec(const or_csr *a, const double *x, double *out);

/* ScaledProblem::new. Returns a heap object; free with or_scaled_free. */
or_scaled *or_scale(const or_lp *p, const or_scaling_params *params);
void or_scaled_free(or_scaled *s);
/* Builds an or_scaled from an already scaled problem (for method tests). */
or_scaled *or_scaled_from_arrays(const or_csr *a, const double *c, const double *lc, const double *uc,
                                 const double *lv, const double *uv, const double *row_scale,
                                 const double *col_scale, double bound_scale, double obj_scale);
double or_finite_bound_norm(const double *lo, const double *hi, int64_t m);
double or_norm2(const double *v, int64_t n);

/* The CPU backend's methods on a scaled problem (cpu/src/lib.rs:137-259). */
void or_spmv(const or_scaled *p, const double *x, double *out);
void or_spmv_t(const or_scaled *p, const double *y, double *out);
void or_primal_step(const or_scaled *p, const double *x, const double *aty, double tau, double *x_t, double *x_bar);
void or_dual_step(const or_scaled *p, const double *y, const double *ax_bar, double sigma, double *y_t);
void or_halpern_step(double *z, const double *z_t, const double *anchor, double w, double rho, int64_t n);
double or_distance_sq(const double *a, const double *b, int64_t n);
void or_sub(const double *a, const double *b, double *out, int64_t n);
double or_norm_sq(const double *v, int64_t n);
double or_dot(const double *a, const double *b, int64_t n);
void or_scale_vec(double *v, double alpha, int64_t n);
or_primal_stats or_primal_stats_of(const or_scaled *p, const double *x, const double *ax);
or_dual_stats or_dual_stats_of(const or_scaled *p, const double *y, const double *aty);
double or_step_dual_residual_sq(const or_scaled *p, const double *x, const double *x_t, const double *aty,
                                const double *aty_t, double tau);
or_ray_stats or_primal_ray_stats(const or_scaled *p, const double *dx, const double *adx);
or_ray_stats or_dual_ray_stats(const or_scaled *p, const double *dy, const double *atdy);

/* PidState (solver.rs:445-480), exposed for its unit tests. */
typedef struct {
    double integral;
    int has_prev;
    double prev_error;
    double best_weight;
    double best_balance;
} or_pid;
void or_pid_init(or_pid *pid, double omega);
double or_pid_update(or_pid *pid, const or_solver_params *p, double omega, double dx, double dy, double rel_primal,
                     double rel_dual);

/* solve() (solver.rs:154-157). x_out (n) and y_out (m) may be NULL. */
int or_solve(const or_lp *p, const or_solver_params *params, or_solution *sol, double *x_out, double *y_out);

/* Runs exactly `iterations` PDHG iterations of Solver::run on an already
 * scaled problem and returns the elapsed seconds (a bounded CPU-baseline
 * sample; the check work every check_every iterations is included). */
double or_time_iterations(const or_lp *p, const or_solver_params *params, uint64_t iterations,
                          double *scale_seconds);

/* check() (check/src/lib.rs:62-118). */
void or_check(const or_lp *p, const double *x, const double *y, or_kkt_report *out);

#ifdef __cplusplus
}
#endif
#endif
