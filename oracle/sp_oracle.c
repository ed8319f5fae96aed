/*
 * Oracle: CPU restatement of saddlepoint's reference path (test infrastructure
 * only; see sp_oracle.h for who may load it). Each function cites the
 * reference file:line it restates. Compiled with -ffp-contract=off so that no
 * product or sum is fused, as in the Rust reference.
 */
#include "sp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

void or_set_threads(int threads) { g_threads = threads < 1 ? 1 : threads; }
int or_get_threads(void) { return g_threads; }

static int par(int64_t n) { return g_threads > 1 && n >= OR_PAR_THRESHOLD; }

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- reductions: cpu/src/lib.rs:37-56 ------------------------------------ */

/* A reduction term is a function of the element index; the closures of the
 * Rust code become small inline functions that take a context pointer. */
typedef void (*term_fn)(const void *ctx, int64_t i, double *out);

/* Sums K terms over [0, n): serial left fold below the threshold
 * (lib.rs:38-40, 46-47), else per-thread partials combined in thread order. */
static void sumk(int64_t n, int k, term_fn f, const void *ctx, double *acc) {
    for (int j = 0; j < k; ++j) acc[j] = 0.0;
    if (!par(n)) {
        double t[4];
        for (int64_t i = 0; i < n; ++i) {
            f(ctx, i, t);
            for (int j = 0; j < k; ++j) acc[j] += t[j];
        }
        return;
    }
    int nt = g_threads;
    double *part = calloc((size_t)nt * 4, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int id = omp_get_thread_num();
        int cnt = omp_get_num_threads();
#else
        int id = 0, cnt = 1;
#endif
        int64_t lo = n * id / cnt, hi = n * (id + 1) / cnt;
        double s[4] = {0, 0, 0, 0}, t[4];
        for (int64_t i = lo; i < hi; ++i) {
            f(ctx, i, t);
            for (int j = 0; j < k; ++j) s[j] += t[j];
        }
        for (int j = 0; j < k; ++j) part[id * 4 + j] = s[j];
    }
    for (int id = 0; id < nt; ++id)
        for (int j = 0; j < k; ++j) acc[j] += part[id * 4 + j];
    free(part);
}

/* ---- CSR: problem.rs:88-126 ---------------------------------------------- */

void or_csr_transpose(const or_csr *a, int64_t *counts, int64_t *col_idx, double *values) {
    /* problem.rs:88-109: counting sort by column; rows come out sorted. */
    memset(counts, 0, sizeof(int64_t) * (size_t)(a->ncols + 1));
    for (int64_t k = 0; k < a->nnz; ++k) counts[a->col_idx[k] + 1] += 1;
    for (int64_t j = 0; j < a->ncols; ++j) counts[j + 1] += counts[j];
    int64_t *next = malloc(sizeof(int64_t) * (size_t)(a->ncols + 1));
    memcpy(next, counts, sizeof(int64_t) * (size_t)(a->ncols + 1));
    for (int64_t i = 0; i < a->nrows; ++i) {
        for (int64_t k = a->row_ptr[i]; k < a->row_ptr[i + 1]; ++k) {
            int64_t c = a->col_idx[k];
            int64_t q = next[c]++;
            col_idx[q] = i;
            values[q] = a->values[k];
        }
    }
    free(next);
}

static double row_dot(const or_csr *a, int64_t i, const double *x) {
    /* problem.rs:112-117 / cpu lib.rs:59-62: each row summed left to right. */
    double s = 0.0;
    for (int64_t k = a->row_ptr[i]; k < a->row_ptr[i + 1]; ++k) s += a->values[k] * x[a->col_idx[k]];
    return s;
}

void or_csr_mul_vec(const or_csr *a, const double *x, double *out) {
    for (int64_t i = 0; i < a->nrows; ++i) out[i] = row_dot(a, i, x);
}

/* csr_mul (cpu/src/lib.rs:58-68): rows in parallel above the threshold; each
 * row's sum is serial either way, so the result does not depend on threads. */
static void csr_mul(const or_csr *a, const double *x, double *out) {
    if (!par(a->nnz)) {
        or_csr_mul_vec(a, x, out);
        return;
    }
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (int64_t i = 0; i < a->nrows; ++i) out[i] = row_dot(a, i, x);
}

static void csr_scale(or_csr *a, const double *row, const double *col) {
    /* problem.rs:120-126: v *= row_i * col_j (the product is formed first).
     * Every element is its own product, so threads change nothing. */
#pragma omp parallel for schedule(dynamic, 1024) num_threads(g_threads) if (par(a->nnz))
    for (int64_t i = 0; i < a->nrows; ++i) {
        double ri = row[i];
        for (int64_t k = a->row_ptr[i]; k < a->row_ptr[i + 1]; ++k) a->values[k] *= ri * col[a->col_idx[k]];
    }
}

static void csr_clone(const or_csr *src, or_csr *dst) {
    dst->nrows = src->nrows;
    dst->ncols = src->ncols;
    dst->nnz = src->nnz;
    dst->row_ptr = malloc(sizeof(int64_t) * (size_t)(src->nrows + 1));
    dst->col_idx = malloc(sizeof(int64_t) * (size_t)(src->nnz ? src->nnz : 1));
    dst->values = malloc(sizeof(double) * (size_t)(src->nnz ? src->nnz : 1));
    memcpy(dst->row_ptr, src->row_ptr, sizeof(int64_t) * (size_t)(src->nrows + 1));
    if (src->nnz) {
        memcpy(dst->col_idx, src->col_idx, sizeof(int64_t) * (size_t)src->nnz);
        memcpy(dst->values, src->values, sizeof(double) * (size_t)src->nnz);
    }
}

static void csr_transpose_alloc(const or_csr *a, or_csr *t) {
    t->nrows = a->ncols;
    t->ncols = a->nrows;
    t->nnz = a->nnz;
    t->row_ptr = malloc(sizeof(int64_t) * (size_t)(a->ncols + 1));
    t->col_idx = malloc(sizeof(int64_t) * (size_t)(a->nnz ? a->nnz : 1));
    t->values = malloc(sizeof(double) * (size_t)(a->nnz ? a->nnz : 1));
    or_csr_transpose(a, t->row_ptr, t->col_idx, t->values);
}

static void csr_free(or_csr *a) {
    free(a->row_ptr);
    free(a->col_idx);
    free(a->values);
}

/* ---- scaling: scaling.rs:53-201 ------------------------------------------ */

void or_default_scaling_params(or_scaling_params *p) {
    /* scaling.rs:25-34 */
    p->geometric_mean_passes = 12;
    p->ruiz_iterations = 10;
    p->has_pock_chambolle = 1;
    p->pock_chambolle_alpha = 1.0;
    p->bound_objective_rescaling = 1;
}

static double inv_sqrt(double v) { return (v > 0.0 && isfinite(v)) ? 1.0 / sqrt(v) : 1.0; } /* :139-141 */

static double fmax_r(double a, double b) { return a > b ? a : b; } /* f64::max for non-NaN */
static double fmin_r(double a, double b) { return a < b ? a : b; }

/* Column-major positions of a's nonzeros, each column's in row order (the
 * order of the counting sort, problem.rs:88-109). A column fold that walks
 * them visits its terms in the order the serial row-major loop of the
 * reference adds them, so the parallel passes below give the serial bits. */
typedef struct {
    int64_t *ptr, *pos;
} colmap;

static int colmap_build(const or_csr *a, colmap *cm) {
    cm->ptr = calloc((size_t)(a->ncols + 1), sizeof(int64_t));
    cm->pos = malloc(sizeof(int64_t) * (size_t)(a->nnz ? a->nnz : 1));
    int64_t *next = malloc(sizeof(int64_t) * (size_t)(a->ncols + 1));
    if (!cm->ptr || !cm->pos || !next) {
        free(cm->ptr), free(cm->pos), free(next);
        cm->ptr = cm->pos = NULL;
        return 0;
    }
    for (int64_t k = 0; k < a->nnz; ++k) cm->ptr[a->col_idx[k] + 1] += 1;
    for (int64_t j = 0; j < a->ncols; ++j) cm->ptr[j + 1] += cm->ptr[j];
    memcpy(next, cm->ptr, sizeof(int64_t) * (size_t)(a->ncols + 1));
    for (int64_t k = 0; k < a->nnz; ++k) cm->pos[next[a->col_idx[k]]++] = k;
    free(next);
    return 1;
}

static void colmap_free(colmap *cm) {
    free(cm->ptr), free(cm->pos);
}

/* min_max (scaling.rs:159-178). With a column map (large matrices, several
 * threads) rows and columns are folded in parallel, each in the serial order. */
static void min_max(const or_csr *a, const colmap *cm, double *rmin, double *rmax, double *cmin, double *cmax) {
    if (cm) {
#pragma omp parallel for schedule(dynamic, 1024) num_threads(g_threads)
        for (int64_t i = 0; i < a->nrows; ++i) {
            double lo = INFINITY, hi = 0.0;
            for (int64_t k = a->row_ptr[i]; k < a->row_ptr[i + 1]; ++k) {
                double v = fabs(a->values[k]);
                lo = fmin_r(lo, v);
                hi = fmax_r(hi, v);
            }
            rmin[i] = hi == 0.0 ? 0.0 : lo;
            rmax[i] = hi;
        }
#pragma omp parallel for schedule(dynamic, 1024) num_threads(g_threads)
        for (int64_t j = 0; j < a->ncols; ++j) {
            double lo = INFINITY, hi = 0.0;
            for (int64_t q = cm->ptr[j]; q < cm->ptr[j + 1]; ++q) {
                double v = fabs(a->values[cm->pos[q]]);
                lo = fmin_r(lo, v);
                hi = fmax_r(hi, v);
            }
            cmin[j] = hi == 0.0 ? 0.0 : lo;
            cmax[j] = hi;
        }
        return;
    }
    for (int64_t i = 0; i < a->nrows; ++i) rmin[i] = INFINITY, rmax[i] = 0.0;
    for (int64_t j = 0; j < a->ncols; ++j) cmin[j] = INFINITY, cmax[j] = 0.0;
    for (int64_t i = 0; i < a->nrows; ++i) {
        for (int64_t k = a->row_ptr[i]; k < a->row_ptr[i + 1]; ++k) {
            int64_t j = a->col_idx[k];
            double v = fabs(a->values[k]);
            rmin[i] = fmin_r(rmin[i], v);
            rmax[i] = fmax_r(rmax[i], v);
            cmin[j] = fmin_r(cmin[j], v);
            cmax[j] = fmax_r(cmax[j], v);
        }
    }
    for (int64_t i = 0; i < a->nrows; ++i)
        if (rmax[i] == 0.0) rmin[i] = 0.0;
    for (int64_t j = 0; j < a->ncols; ++j)
        if (cmax[j] == 0.0) cmin[j] = 0.0;
}

/* abs_stats (scaling.rs:144-155) with the three folds the code uses. A sum
 * adds a line's terms left to right (a column's in row order), also in the
 * parallel form. */
enum { FOLD_MAX, FOLD_POW };
static double fold_term(int fold, double e, double acc, double v) {
    return fold == FOLD_MAX ? fmax_r(acc, v) : acc + pow(v, e);
}

static void abs_stats(const or_csr *a, const colmap *cm, int fold, double e, double *rows, double *cols) {
    if (cm) {
#pragma omp parallel for schedule(dynamic, 1024) num_threads(g_threads)
        for (int64_t i = 0; i < a->nrows; ++i) {
            double s = 0.0;
            for (int64_t k = a->row_ptr[i]; k < a->row_ptr[i + 1]; ++k) s = fold_term(fold, e, s, fabs(a->values[k]));
            rows[i] = s;
        }
#pragma omp parallel for schedule(dynamic, 1024) num_threads(g_threads)
        for (int64_t j = 0; j < a->ncols; ++j) {
            double s = 0.0;
            for (int64_t q = cm->ptr[j]; q < cm->ptr[j + 1]; ++q) s = fold_term(fold, e, s, fabs(a->values[cm->pos[q]]));
            cols[j] = s;
        }
        return;
    }
    for (int64_t i = 0; i < a->nrows; ++i) rows[i] = 0.0;
    for (int64_t j = 0; j < a->ncols; ++j) cols[j] = 0.0;
    for (int64_t i = 0; i < a->nrows; ++i) {
        for (int64_t k = a->row_ptr[i]; k < a->row_ptr[i + 1]; ++k) {
            int64_t j = a->col_idx[k];
            double v = fabs(a->values[k]);
            if (fold == FOLD_MAX) {
                rows[i] = fmax_r(rows[i], v);
                cols[j] = fmax_r(cols[j], v);
            } else {
                double t = pow(v, e);
                rows[i] = rows[i] + t;
                cols[j] = cols[j] + t;
            }
        }
    }
}

static void apply(or_csr *a, double *row_scale, double *col_scale, const double *r, const double *c) {
    /* scaling.rs:180-184 */
    csr_scale(a, r, c);
    for (int64_t i = 0; i < a->nrows; ++i) row_scale[i] *= r[i];
    for (int64_t j = 0; j < a->ncols; ++j) col_scale[j] *= c[j];
}

double or_norm2(const double *v, int64_t n) {
    /* scaling.rs:186-188 */
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += v[i] * v[i];
    return sqrt(s);
}

double or_finite_bound_norm(const double *lo, const double *hi, int64_t m) {
    /* scaling.rs:191-201: infinite bounds count as 0, an equality row once. */
    double s = 0.0;
    for (int64_t i = 0; i < m; ++i) {
        double l = isfinite(lo[i]) ? lo[i] : 0.0;
        double u = isfinite(hi[i]) ? hi[i] : 0.0;
        s += (l == u) ? l * l : l * l + u * u;
    }
    return sqrt(s);
}

static double *dup(const double *v, int64_t n) {
    double *o = malloc(sizeof(double) * (size_t)(n ? n : 1));
    if (n) memcpy(o, v, sizeof(double) * (size_t)n);
    return o;
}

static void compute_factors(or_scaled *s) {
    /* scaling.rs:124-131 */
    int64_t m = s->a.nrows, n = s->a.ncols;
    s->primal_factor = malloc(sizeof(double) * (size_t)(m ? m : 1));
    s->dual_factor = malloc(sizeof(double) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < m; ++i) s->primal_factor[i] = 1.0 / (s->row_scale[i] * s->bound_scale);
    for (int64_t j = 0; j < n; ++j) s->dual_factor[j] = 1.0 / (s->col_scale[j] * s->obj_scale);
}

or_scaled *or_scale(const or_lp *p, const or_scaling_params *params) {
    /* ScaledProblem::new, scaling.rs:53-103 */
    int64_t m = p->a.nrows, n = p->a.ncols;
    size_t mb = sizeof(double) * (size_t)(m ? m : 1), nb = sizeof(double) * (size_t)(n ? n : 1);
    or_scaled *s = calloc(1, sizeof(or_scaled));
    csr_clone(&p->a, &s->a);
    s->row_scale = malloc(mb);
    s->col_scale = malloc(nb);
    double *ones_m = malloc(mb), *ones_n = malloc(nb);
    for (int64_t i = 0; i < m; ++i) s->row_scale[i] = 1.0, ones_m[i] = 1.0;
    for (int64_t j = 0; j < n; ++j) s->col_scale[j] = 1.0, ones_n[j] = 1.0;
    double *rmin = malloc(mb), *rmax = malloc(mb), *cmin = malloc(nb), *cmax = malloc(nb);
    double *r = malloc(mb), *c = malloc(nb);
    colmap cmap = {NULL, NULL};
    const colmap *cm = par(s->a.nnz) && colmap_build(&s->a, &cmap) ? &cmap : NULL;

    for (int64_t pass = 0; pass < params->geometric_mean_passes; ++pass) {
        min_max(&s->a, cm, rmin, rmax, cmin, cmax);
        for (int64_t i = 0; i < m; ++i) r[i] = inv_sqrt(rmin[i] * rmax[i]);
        apply(&s->a, s->row_scale, s->col_scale, r, ones_n);
        min_max(&s->a, cm, rmin, rmax, cmin, cmax);
        for (int64_t j = 0; j < n; ++j) c[j] = inv_sqrt(cmin[j] * cmax[j]);
        apply(&s->a, s->row_scale, s->col_scale, ones_m, c);
    }
    for (int64_t it = 0; it < params->ruiz_iterations; ++it) {
        abs_stats(&s->a, cm, FOLD_MAX, 0.0, rmax, cmax);
        for (int64_t i = 0; i < m; ++i) r[i] = inv_sqrt(rmax[i]);
        for (int64_t j = 0; j < n; ++j) c[j] = inv_sqrt(cmax[j]);
        apply(&s->a, s->row_scale, s->col_scale, r, c);
    }
    if (params->has_pock_chambolle) {
        double alpha = params->pock_chambolle_alpha;
        abs_stats(&s->a, cm, FOLD_POW, alpha, rmax, cmax); /* rows use |a|^α */
        for (int64_t i = 0; i < m; ++i) r[i] = inv_sqrt(rmax[i]);
        abs_stats(&s->a, cm, FOLD_POW, 2.0 - alpha, rmin, cmax); /* columns use |a|^(2-α) */
        for (int64_t j = 0; j < n; ++j) c[j] = inv_sqrt(cmax[j]);
        apply(&s->a, s->row_scale, s->col_scale, r, c);
    }

    s->c = malloc(nb);
    s->lv = malloc(nb);
    s->uv = malloc(nb);
    s->lc = malloc(mb);
    s->uc = malloc(mb);
    for (int64_t j = 0; j < n; ++j) {
        s->c[j] = p->c[j] * s->col_scale[j];
        s->lv[j] = p->lv[j] / s->col_scale[j];
        s->uv[j] = p->uv[j] / s->col_scale[j];
    }
    for (int64_t i = 0; i < m; ++i) {
        s->lc[i] = p->lc[i] * s->row_scale[i];
        s->uc[i] = p->uc[i] * s->row_scale[i];
    }
    s->bound_scale = 1.0;
    s->obj_scale = 1.0;
    if (params->bound_objective_rescaling) {
        s->bound_scale = 1.0 / (1.0 + or_finite_bound_norm(s->lc, s->uc, m));
        s->obj_scale = 1.0 / (1.0 + or_norm2(s->c, n));
        for (int64_t i = 0; i < m; ++i) s->lc[i] *= s->bound_scale;
        for (int64_t i = 0; i < m; ++i) s->uc[i] *= s->bound_scale;
        for (int64_t j = 0; j < n; ++j) s->lv[j] *= s->bound_scale;
        for (int64_t j = 0; j < n; ++j) s->uv[j] *= s->bound_scale;
        for (int64_t j = 0; j < n; ++j) s->c[j] *= s->obj_scale;
    }
    csr_transpose_alloc(&s->a, &s->at);
    compute_factors(s);
    colmap_free(&cmap);
    free(ones_m), free(ones_n), free(rmin), free(rmax), free(cmin), free(cmax), free(r), free(c);
    return s;
}

or_scaled *or_scaled_from_arrays(const or_csr *a, const double *c, const double *lc, const double *uc,
                                 const double *lv, const double *uv, const double *row_scale,
                                 const double *col_scale, double bound_scale, double obj_scale) {
    or_scaled *s = calloc(1, sizeof(or_scaled));
    int64_t m = a->nrows, n = a->ncols;
    csr_clone(a, &s->a);
    csr_transpose_alloc(&s->a, &s->at);
    s->c = dup(c, n);
    s->lv = dup(lv, n);
    s->uv = dup(uv, n);
    s->lc = dup(lc, m);
    s->uc = dup(uc, m);
    s->row_scale = dup(row_scale, m);
    s->col_scale = dup(col_scale, n);
    s->bound_scale = bound_scale;
    s->obj_scale = obj_scale;
    compute_factors(s);
    return s;
}

void or_scaled_free(or_scaled *s) {
    if (!s) return;
    csr_free(&s->a);
    csr_free(&s->at);
    free(s->c), free(s->lc), free(s->uc), free(s->lv), free(s->uv);
    free(s->row_scale), free(s->col_scale), free(s->primal_factor), free(s->dual_factor);
    free(s);
}

/* ---- CPU backend: cpu/src/lib.rs:70-269 ---------------------------------- */

static inline double clampd(double v, double lo, double hi) {
    /* Rust f64::clamp: NaN passes through. */
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

static inline double excess(double v, double lo, double hi) { /* lib.rs:72-80 */
    return v < lo ? v - lo : (v > hi ? v - hi : 0.0);
}

static inline double bound_term(double v, double lo, double hi) { /* lib.rs:84-92 */
    if (v > 0.0 && isfinite(lo)) return v * lo;
    if (v < 0.0 && isfinite(hi)) return v * hi;
    return 0.0;
}

static inline double forbidden(double v, double lo, double hi) { /* lib.rs:97-99 */
    return ((v > 0.0 && lo == -INFINITY) || (v < 0.0 && hi == INFINITY)) ? v : 0.0;
}

static inline double recession_violation(double v, double lo, double hi) { /* lib.rs:265-269 */
    double l = isfinite(lo) ? 0.0 : -INFINITY;
    double h = isfinite(hi) ? 0.0 : INFINITY;
    return excess(v, l, h);
}

void or_spmv(const or_scaled *p, const double *x, double *out) { csr_mul(&p->a, x, out); }
void or_spmv_t(const or_scaled *p, const double *y, double *out) { csr_mul(&p->at, y, out); }

void or_primal_step(const or_scaled *p, const double *x, const double *aty, double tau, double *x_t, double *x_bar) {
    /* lib.rs:147-160: two passes, as in the reference. */
    int64_t n = p->a.ncols;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (par(n))
    for (int64_t j = 0; j < n; ++j) x_t[j] = clampd(x[j] - tau * (p->c[j] - aty[j]), p->lv[j], p->uv[j]);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (par(n))
    for (int64_t j = 0; j < n; ++j) x_bar[j] = 2.0 * x_t[j] - x[j];
}

void or_dual_step(const or_scaled *p, const double *y, const double *ax_bar, double sigma, double *y_t) {
    /* lib.rs:162-175 */
    int64_t m = p->a.nrows;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (par(m))
    for (int64_t i = 0; i < m; ++i) {
        double v = ax_bar[i] - y[i] / sigma;
        y_t[i] = y[i] - sigma * (ax_bar[i] - clampd(v, p->lc[i], p->uc[i]));
    }
}

void or_halpern_step(double *z, const double *z_t, const double *anchor, double w, double rho, int64_t n) {
    /* lib.rs:177-187 */
#pragma omp parallel for schedule(static) num_threads(g_threads) if (par(n))
    for (int64_t i = 0; i < n; ++i) z[i] = w * ((1.0 + rho) * z_t[i] - rho * z[i]) + (1.0 - w) * anchor[i];
}

void or_sub(const double *a, const double *b, double *out, int64_t n) {
#pragma omp parallel for schedule(static) num_threads(g_threads) if (par(n))
    for (int64_t i = 0; i < n; ++i) out[i] = a[i] - b[i];
}

void or_scale_vec(double *v, double alpha, int64_t n) {
#pragma omp parallel for schedule(static) num_threads(g_threads) if (par(n))
    for (int64_t i = 0; i < n; ++i) v[i] *= alpha;
}

typedef struct {
    const double *a, *b, *c, *d, *e, *f, *g;
    double s;
} ctx7;

static void t_dist(const void *v, int64_t i, double *o) {
    const ctx7 *x = v;
    double d = x->a[i] - x->b[i];
    o[0] = d * d;
}
static void t_dot(const void *v, int64_t i, double *o) {
    const ctx7 *x = v;
    o[0] = x->a[i] * x->b[i];
}

double or_distance_sq(const double *a, const double *b, int64_t n) { /* lib.rs:189-191 */
    ctx7 c = {a, b};
    double r;
    sumk(n, 1, t_dist, &c, &r);
    return r;
}
double or_norm_sq(const double *v, int64_t n) { return or_dot(v, v, n); } /* lib.rs:198-200 */
double or_dot(const double *a, const double *b, int64_t n) {             /* lib.rs:202-204 */
    ctx7 c = {a, b};
    double r;
    sumk(n, 1, t_dot, &c, &r);
    return r;
}

/* primal_stats, lib.rs:211-215 */
static void t_pres(const void *v, int64_t i, double *o) {
    const ctx7 *x = v; /* a=ax b=lc c=uc d=pf */
    double e = excess(x->a[i], x->b[i], x->c[i]) * x->d[i];
    o[0] = e * e;
}
or_primal_stats or_primal_stats_of(const or_scaled *p, const double *x, const double *ax) {
    or_primal_stats r;
    ctx7 c1 = {ax, p->lc, p->uc, p->primal_factor};
    sumk(p->a.nrows, 1, t_pres, &c1, &r.residual_sq);
    ctx7 c2 = {p->c, x};
    sumk(p->a.ncols, 1, t_dot, &c2, &r.objective);
    return r;
}

/* dual_stats, lib.rs:217-224 */
static void t_drows(const void *v, int64_t i, double *o) {
    const ctx7 *x = v; /* a=y b=lc c=uc */
    o[0] = bound_term(x->a[i], x->b[i], x->c[i]);
}
static void t_dcols(const void *v, int64_t j, double *o) {
    const ctx7 *x = v; /* a=c b=aty c=lv d=uv e=df */
    double r = x->a[j] - x->b[j];
    double f = forbidden(r, x->c[j], x->d[j]) * x->e[j];
    o[0] = f * f;
    o[1] = bound_term(r, x->c[j], x->d[j]);
}
or_dual_stats or_dual_stats_of(const or_scaled *p, const double *y, const double *aty) {
    double rows, cols[2];
    ctx7 c1 = {y, p->lc, p->uc};
    sumk(p->a.nrows, 1, t_drows, &c1, &rows);
    ctx7 c2 = {p->c, aty, p->lv, p->uv, p->dual_factor};
    sumk(p->a.ncols, 2, t_dcols, &c2, cols);
    or_dual_stats r = {cols[0], rows + cols[1]};
    return r;
}

/* step_dual_residual_sq, lib.rs:226-237 */
static void t_step(const void *v, int64_t j, double *o) {
    const ctx7 *x = v; /* a=aty b=aty_t c=x_t d=x e=df s=inv_tau */
    double r = (x->a[j] - x->b[j] - (x->c[j] - x->d[j]) * x->s) * x->e[j];
    o[0] = r * r;
}
double or_step_dual_residual_sq(const or_scaled *p, const double *x, const double *x_t, const double *aty,
                                const double *aty_t, double tau) {
    ctx7 c = {aty, aty_t, x_t, x, p->dual_factor, NULL, NULL, 1.0 / tau};
    double r;
    sumk(p->a.ncols, 1, t_step, &c, &r);
    return r;
}

/* primal_ray_stats, lib.rs:239-247 */
static void t_prow(const void *v, int64_t i, double *o) {
    const ctx7 *x = v; /* a=adx b=lc c=uc */
    double r = recession_violation(x->a[i], x->b[i], x->c[i]);
    o[0] = r * r;
}
static void t_pcol(const void *v, int64_t j, double *o) {
    const ctx7 *x = v; /* a=dx b=lv c=uv */
    double r = recession_violation(x->a[j], x->b[j], x->c[j]);
    o[0] = r * r;
    o[1] = x->a[j] * x->a[j];
}
or_ray_stats or_primal_ray_stats(const or_scaled *p, const double *dx, const double *adx) {
    double rows, cols[2], obj;
    ctx7 c1 = {adx, p->lc, p->uc};
    sumk(p->a.nrows, 1, t_prow, &c1, &rows);
    ctx7 c2 = {dx, p->lv, p->uv};
    sumk(p->a.ncols, 2, t_pcol, &c2, cols);
    ctx7 c3 = {p->c, dx};
    sumk(p->a.ncols, 1, t_dot, &c3, &obj);
    or_ray_stats r = {obj, rows + cols[0], cols[1]};
    return r;
}

/* dual_ray_stats, lib.rs:249-259 */
static void t_drrow(const void *v, int64_t i, double *o) {
    const ctx7 *x = v; /* a=dy b=lc c=uc */
    double f = forbidden(x->a[i], x->b[i], x->c[i]);
    o[0] = f * f;
    o[1] = bound_term(x->a[i], x->b[i], x->c[i]);
}
static void t_drcol(const void *v, int64_t j, double *o) {
    const ctx7 *x = v; /* a=atdy b=lv c=uv */
    double r = -x->a[j];
    double f = forbidden(r, x->b[j], x->c[j]);
    o[0] = f * f;
    o[1] = bound_term(r, x->b[j], x->c[j]);
}
or_ray_stats or_dual_ray_stats(const or_scaled *p, const double *dy, const double *atdy) {
    double rows[2], cols[2], norm;
    ctx7 c1 = {dy, p->lc, p->uc};
    sumk(p->a.nrows, 2, t_drrow, &c1, rows);
    ctx7 c2 = {atdy, p->lv, p->uv};
    sumk(p->a.ncols, 2, t_drcol, &c2, cols);
    ctx7 c3 = {dy, dy};
    sumk(p->a.nrows, 1, t_dot, &c3, &norm);
    or_ray_stats r = {rows[1] + cols[1], rows[0] + cols[0], norm};
    return r;
}

/* ---- solver: solver.rs:70-480 -------------------------------------------- */

void or_default_solver_params(or_solver_params *p) {
    /* solver.rs:70-93 */
    p->tolerance = 1e-6;
    p->infeasibility_tolerance = 1e-8;
    p->max_iterations = 1000000;
    p->time_limit_s = -1.0;
    p->check_every = 64;
    p->step_size_factor = 0.998;
    p->reflection = 1.0;
    p->restart_sufficient = 0.2;
    p->restart_necessary = 0.5;
    p->restart_artificial = 0.36;
    p->pid_kp = 0.99;
    p->pid_ki = 0.01;
    p->pid_kd = 0.0;
    p->pid_discount = 0.3;
    p->log_every_checks = 0;
    p->power_iterations = 200;
    or_default_scaling_params(&p->scaling);
    p->termination = 0;
}

void or_pid_init(or_pid *pid, double omega) { /* solver.rs:453-455 */
    pid->integral = 0.0;
    pid->has_prev = 0;
    pid->prev_error = 0.0;
    pid->best_weight = omega;
    pid->best_balance = INFINITY;
}

double or_pid_update(or_pid *pid, const or_solver_params *p, double omega, double dx, double dy, double rel_primal,
                     double rel_dual) {
    /* solver.rs:459-479 */
#define SANE(v) ((v) >= 1e-16 && (v) <= 1e12)
    double ratio = rel_dual / rel_primal;
    if (!SANE(dx) || !SANE(dy) || !(ratio >= 1e-8 && ratio <= 1e8)) {
        pid->integral = 0.0;
        pid->has_prev = 0;
        return pid->best_weight;
    }
#undef SANE
    double e = log(omega * dx / dy);
    pid->integral = p->pid_discount * pid->integral + e;
    double derivative = pid->has_prev ? e - pid->prev_error : 0.0;
    pid->has_prev = 1;
    pid->prev_error = e;
    double next = exp(log(omega) - (p->pid_kp * e + p->pid_ki * pid->integral + p->pid_kd * derivative));
    double balance = fabs(log10(ratio));
    if (balance < pid->best_balance) {
        pid->best_balance = balance;
        pid->best_weight = next;
    }
    return next;
}

typedef struct {
    double primal_obj, dual_obj, primal_res, dual_res;
} kkt_t;

static double kkt_gap(const kkt_t *k) { return fabs(k->primal_obj - k->dual_obj); }
static int kkt_converged(const kkt_t *k, double eps, double b_norm, double c_norm) { /* solver.rs:147-151 */
    return k->primal_res <= eps * (1.0 + b_norm) && k->dual_res <= eps * (1.0 + c_norm) &&
           kkt_gap(k) <= eps * (1.0 + fabs(k->primal_obj) + fabs(k->dual_obj));
}

typedef struct {
    const or_lp *problem;
    const or_solver_params *params;
    or_scaled *s;
    int64_t m, n;
    double *x, *y, *x_t, *y_t, *x_bar, *x_anchor, *y_anchor, *ax, *aty, *tmp_n, *tmp_m, *tmp_m2, *c;
} solver_t;

static int estimate_norm(solver_t *S, double *out) {
    /* solver.rs:215-242 */
    int64_t n = S->n;
    if (n == 0 || S->s->a.nnz == 0) {
        *out = 1.0;
        return OR_OK;
    }
    double *v = malloc(sizeof(double) * (size_t)n);
    for (int64_t j = 0; j < n; ++j) v[j] = 1.0 + 0.01 * (double)(j % 7);
    double s = 1.0 / sqrt(or_norm_sq(v, n));
    or_scale_vec(v, s, n);
    double lambda = 0.0;
    for (int64_t it = 0; it < S->params->power_iterations; ++it) {
        or_spmv(S->s, v, S->tmp_m);
        or_spmv_t(S->s, S->tmp_m, v);
        double norm = sqrt(or_norm_sq(v, n));
        if (isnan(norm) || norm <= 0.0 || isinf(norm)) {
            free(v);
            return OR_ERR_NUMERICAL;
        }
        or_scale_vec(v, 1.0 / norm, n);
        int done = fabs(norm - lambda) <= 1e-6 * norm;
        lambda = norm;
        if (done) break;
    }
    free(v);
    *out = sqrt(lambda) * 1.01;
    return OR_OK;
}

static double objective_factor(const or_scaled *s) { return 1.0 / (s->bound_scale * s->obj_scale); }

static kkt_t kkt(solver_t *S) {
    /* solver.rs:245-258 */
    double f = objective_factor(S->s);
    or_spmv(S->s, S->x_t, S->tmp_m);
    or_primal_stats ps = or_primal_stats_of(S->s, S->x_t, S->tmp_m);
    or_spmv_t(S->s, S->y_t, S->tmp_n);
    or_dual_stats ds = or_dual_stats_of(S->s, S->y_t, S->tmp_n);
    kkt_t k = {ps.objective * f + S->problem->c0, ds.objective * f + S->problem->c0, sqrt(ps.residual_sq),
               sqrt(ds.residual_sq)};
    return k;
}

static double step_dual_objective(solver_t *S, double tau) {
    /* solver.rs:263-271 */
    double rows = or_dual_stats_of(S->s, S->y_t, S->c).objective;
    int64_t n = S->n;
    double lx = or_dot(S->c, S->x_t, n) - or_dot(S->aty, S->x_t, n) +
                (or_dot(S->x_t, S->x_t, n) - or_dot(S->x, S->x_t, n)) / tau;
    return (rows + lx) * objective_factor(S->s) + S->problem->c0;
}

static int fixed_point_residual(solver_t *S, double eta, double omega, double *out) {
    /* solver.rs:275-288 */
    or_sub(S->x_t, S->x, S->tmp_n, S->n);
    or_sub(S->y_t, S->y, S->tmp_m2, S->m);
    or_spmv(S->s, S->tmp_n, S->tmp_m);
    double dx2 = or_norm_sq(S->tmp_n, S->n);
    double dy2 = or_norm_sq(S->tmp_m2, S->m);
    double cross = or_dot(S->tmp_m2, S->tmp_m, S->m);
    double r2 = omega / eta * dx2 + dy2 / (eta * omega) + 2.0 * cross;
    if (!isfinite(r2)) return OR_ERR_NUMERICAL;
    *out = sqrt(r2 > 0.0 ? r2 : 0.0);
    return OR_OK;
}

static int infeasibility(solver_t *S) {
    /* solver.rs:291-309: returns a status or -1 */
    double eps = S->params->infeasibility_tolerance;
    or_sub(S->y_t, S->y_anchor, S->tmp_m, S->m);
    or_spmv_t(S->s, S->tmp_m, S->tmp_n);
    or_ray_stats d = or_dual_ray_stats(S->s, S->tmp_m, S->tmp_n);
    if (d.objective > 0.0 && sqrt(d.violation_sq) <= eps * d.objective && d.objective > eps * sqrt(d.norm_sq))
        return OR_PRIMAL_INFEASIBLE;
    or_sub(S->x_t, S->x_anchor, S->tmp_n, S->n);
    or_spmv(S->s, S->tmp_n, S->tmp_m);
    or_ray_stats p = or_primal_ray_stats(S->s, S->tmp_n, S->tmp_m);
    if (p.objective < 0.0 && sqrt(p.violation_sq) <= -eps * p.objective && -p.objective > eps * sqrt(p.norm_sq))
        return OR_DUAL_INFEASIBLE;
    return -1;
}

static double *zeros(int64_t n) { return calloc((size_t)(n ? n : 1), sizeof(double)); }

static int validate(const or_lp *p, char *err) {
    /* problem.rs:166-194 (lengths are fixed by the view). */
    for (int pass = 0; pass < 2; ++pass) {
        const double *lo = pass == 0 ? p->lv : p->lc, *hi = pass == 0 ? p->uv : p->uc;
        int64_t len = pass == 0 ? p->a.ncols : p->a.nrows;
        const char *what = pass == 0 ? "variable" : "constraint";
        for (int64_t i = 0; i < len; ++i) {
            if (isnan(lo[i]) || isnan(hi[i])) {
                snprintf(err, 128, "%s %lld is NaN", what, (long long)i);
                return OR_ERR_PROBLEM;
            }
            if (lo[i] > hi[i]) {
                snprintf(err, 128, "%s %lld has lower bound %g above upper bound %g", what, (long long)i, lo[i],
                         hi[i]);
                return OR_ERR_PROBLEM;
            }
        }
    }
    for (int64_t j = 0; j < p->a.ncols; ++j)
        if (isnan(p->c[j])) {
            snprintf(err, 128, "coefficient %lld is NaN", (long long)j);
            return OR_ERR_PROBLEM;
        }
    for (int64_t k = 0; k < p->a.nnz; ++k)
        if (isnan(p->a.values[k])) {
            snprintf(err, 128, "coefficient %lld is NaN", (long long)(p->a.ncols + k));
            return OR_ERR_PROBLEM;
        }
    return OR_OK;
}

static int run(solver_t *S, or_solution *sol, double t_start, double *x_out, double *y_out) {
    /* Solver::run, solver.rs:311-438 */
    const or_solver_params *p = S->params;
    int64_t m = S->m, n = S->n;
    double start = now_s();
    double norm;
    int rc = estimate_norm(S, &norm);
    if (rc) {
        snprintf(sol->error, sizeof sol->error, "numerical breakdown: power iteration collapsed");
        return rc;
    }
    double eta = p->step_size_factor / norm;
    double loop_start = now_s();
    double b_norm = or_finite_bound_norm(S->problem->lc, S->problem->uc, m);
    double c_norm = or_norm2(S->problem->c, n);
    double omega = p->scaling.bound_objective_rescaling ? 1.0 : (1.0 + c_norm) / (1.0 + b_norm);
    or_pid pid;
    or_pid_init(&pid, omega);
    uint64_t total = 0, k = 0, restarts = 0, checks = 0;
    double anchor_residual = INFINITY, prev_residual = INFINITY;
    int status;
    kkt_t kk;
    for (;;) {
        double tau = eta / omega, sigma = eta * omega;
        or_spmv_t(S->s, S->y, S->aty);
        or_primal_step(S->s, S->x, S->aty, tau, S->x_t, S->x_bar);
        or_spmv(S->s, S->x_bar, S->ax);
        or_dual_step(S->s, S->y, S->ax, sigma, S->y_t);
        total += 1;
        int check = (total % p->check_every == 0) || total >= p->max_iterations;
        double residual = 0.0;
        if (check || k == 0) {
            rc = fixed_point_residual(S, eta, omega, &residual);
            if (rc) {
                snprintf(sol->error, sizeof sol->error, "numerical breakdown: iterates are not finite");
                return rc;
            }
        }
        if (k == 0) anchor_residual = residual;
        double step_dual = 0.0, step_obj = 0.0;
        if (check) {
            kk = kkt(S);
            step_dual = sqrt(or_step_dual_residual_sq(S->s, S->x, S->x_t, S->aty, S->tmp_n, tau));
            if (p->termination == 1) step_obj = step_dual_objective(S, tau);
        }
        double w = (double)(k + 1) / (double)(k + 2);
        or_halpern_step(S->x, S->x_t, S->x_anchor, w, p->reflection, n);
        or_halpern_step(S->y, S->y_t, S->y_anchor, w, p->reflection, m);
        k += 1;
        if (!check) continue;
        checks += 1;
        if (p->log_every_checks > 0 && checks % (uint64_t)p->log_every_checks == 0) {
            /* solver.rs:367-378 */
            fprintf(stderr, "%9llu %14.6e %14.6e %9.2e %9.2e %9.2e %9.2e %5llu %8.2fs %7llu %9.2e\n",
                    (unsigned long long)total, kk.primal_obj, kk.dual_obj, kk.primal_res, kk.dual_res,
                    kkt_gap(&kk), omega, (unsigned long long)restarts, now_s() - start,
                    (unsigned long long)k, residual / anchor_residual);
            fflush(stderr);
        }
        kkt_t test = kk;
        if (p->termination == 1) {
            test.dual_obj = step_obj;
            test.dual_res = step_dual;
        }
        if (kkt_converged(&test, p->tolerance, b_norm, c_norm)) {
            status = OR_OPTIMAL;
            break;
        }
        int inf = infeasibility(S);
        if (inf >= 0) {
            status = inf;
            break;
        }
        if (total >= p->max_iterations) {
            status = OR_ITERATION_LIMIT;
            break;
        }
        if (p->time_limit_s >= 0.0 && now_s() - start >= p->time_limit_s) {
            status = OR_TIME_LIMIT;
            break;
        }
        double r0 = anchor_residual;
        int restart = residual <= p->restart_sufficient * r0 ||
                      (residual <= p->restart_necessary * r0 && residual > prev_residual) ||
                      (double)k >= p->restart_artificial * (double)total;
        prev_residual = residual;
        if (restart) {
            double dx = sqrt(or_distance_sq(S->x_t, S->x_anchor, n));
            double dy = sqrt(or_distance_sq(S->y_t, S->y_anchor, m));
            double rel_primal = kk.primal_res / (1.0 + b_norm);
            double rel_dual = step_dual / (1.0 + c_norm);
            omega = or_pid_update(&pid, p, omega, dx, dy, rel_primal, rel_dual);
            memcpy(S->x_anchor, S->x_t, sizeof(double) * (size_t)n);
            memcpy(S->y_anchor, S->y_t, sizeof(double) * (size_t)m);
            memcpy(S->x, S->x_t, sizeof(double) * (size_t)n);
            memcpy(S->y, S->y_t, sizeof(double) * (size_t)m);
            k = 0;
            restarts += 1;
            prev_residual = INFINITY;
        }
    }
    double sign = S->problem->maximize ? -1.0 : 1.0;
    /* solver.rs:422-423: unscale_primal / unscale_dual (scaling.rs:114-121) */
    if (x_out)
        for (int64_t j = 0; j < n; ++j) x_out[j] = S->x_t[j] * S->s->col_scale[j] / S->s->bound_scale;
    if (y_out)
        for (int64_t i = 0; i < m; ++i) y_out[i] = S->y_t[i] * S->s->row_scale[i] / S->s->obj_scale;
    sol->status = status;
    sol->objective = sign * kk.primal_obj;
    sol->dual_objective = sign * kk.dual_obj;
    sol->primal_residual = kk.primal_res;
    sol->dual_residual = kk.dual_res;
    sol->gap = kkt_gap(&kk);
    sol->iterations = total;
    sol->restarts = restarts;
    sol->final_omega = omega;
    double end = now_s();
    sol->solve_time_s = end - start;
    sol->loop_time_s = end - loop_start;
    sol->total_time_s = end - t_start;
    return OR_OK;
}

int or_solve(const or_lp *p, const or_solver_params *params, or_solution *sol, double *x_out, double *y_out) {
    memset(sol, 0, sizeof *sol);
    double t0 = now_s();
    int rc = validate(p, sol->error);
    if (rc) return rc;
    solver_t S = {0};
    S.problem = p;
    S.params = params;
    S.m = p->a.nrows;
    S.n = p->a.ncols;
    S.s = or_scale(p, &params->scaling);
    /* solver.rs:189-205: thirteen vectors, all zero except c. */
    S.x = zeros(S.n), S.y = zeros(S.m), S.x_t = zeros(S.n), S.y_t = zeros(S.m), S.x_bar = zeros(S.n);
    S.x_anchor = zeros(S.n), S.y_anchor = zeros(S.m), S.ax = zeros(S.m), S.aty = zeros(S.n);
    S.tmp_n = zeros(S.n), S.tmp_m = zeros(S.m), S.tmp_m2 = zeros(S.m), S.c = dup(S.s->c, S.n);
    rc = run(&S, sol, t0, x_out, y_out);
    double *vs[] = {S.x, S.y, S.x_t, S.y_t, S.x_bar, S.x_anchor, S.y_anchor, S.ax, S.aty, S.tmp_n, S.tmp_m, S.tmp_m2, S.c};
    for (size_t i = 0; i < sizeof vs / sizeof vs[0]; ++i) free(vs[i]);
    or_scaled_free(S.s);
    return rc;
}

double or_time_iterations(const or_lp *p, const or_solver_params *params, uint64_t iterations,
                          double *scale_seconds) {
    or_solver_params q = *params;
    q.max_iterations = iterations;
    q.tolerance = 0.0;
    or_solution sol;
    double t0 = now_s();
    int rc = or_solve(p, &q, &sol, NULL, NULL);
    (void)t0;
    if (scale_seconds) *scale_seconds = sol.total_time_s - sol.solve_time_s;
    return rc ? -1.0 : sol.loop_time_s;
}

/* ---- checker: check/src/lib.rs:45-118 ------------------------------------ */

static double outside(double v, double lo, double hi) { /* check/lib.rs:49-57 */
    return v < lo ? lo - v : (v > hi ? v - hi : 0.0);
}

void or_check(const or_lp *p, const double *x, const double *y, or_kkt_report *r) {
    int64_t m = p->a.nrows, n = p->a.ncols;
    double *ax = zeros(m), *aty = zeros(n);
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t k = p->a.row_ptr[i]; k < p->a.row_ptr[i + 1]; ++k) {
            int64_t j = p->a.col_idx[k];
            double v = p->a.values[k];
            ax[i] += v * x[j];
            aty[j] += v * y[i];
        }
    }
    double pobj = 0.0;
    for (int64_t j = 0; j < n; ++j) pobj += p->c[j] * x[j];
    pobj += p->c0;
    double cv = 0.0, bv = 0.0;
    for (int64_t i = 0; i < m; ++i) {
        double o = outside(ax[i], p->lc[i], p->uc[i]);
        cv += o * o;
    }
    for (int64_t j = 0; j < n; ++j) {
        double o = outside(x[j], p->lv[j], p->uv[j]);
        bv += o * o;
    }
    cv = sqrt(cv);
    bv = sqrt(bv);
    double dobj = p->c0, dviol = 0.0;
    for (int64_t t = 0; t < m + n; ++t) {
        double v, lo, hi;
        if (t < m) {
            v = y[t], lo = p->lc[t], hi = p->uc[t];
        } else {
            int64_t j = t - m;
            v = p->c[j] - aty[j], lo = p->lv[j], hi = p->uv[j];
        }
        if (v > 0.0) {
            if (isfinite(lo)) dobj += v * lo;
            else dviol += v * v;
        } else if (v < 0.0) {
            if (isfinite(hi)) dobj += v * hi;
            else dviol += v * v;
        }
    }
    dviol = sqrt(dviol);
    double bn = 0.0;
    for (int64_t i = 0; i < m; ++i) {
        double l = isfinite(p->lc[i]) ? p->lc[i] : 0.0;
        double u = (isfinite(p->uc[i]) && p->uc[i] != p->lc[i]) ? p->uc[i] : 0.0;
        bn += l * l;
        bn += u * u;
    }
    bn = sqrt(bn);
    double cn = or_norm2(p->c, n);
    double primal = sqrt(cv * cv + bv * bv);
    r->primal_objective = pobj;
    r->dual_objective = dobj;
    r->constraint_violation = cv;
    r->bound_violation = bv;
    r->dual_violation = dviol;
    r->relative_primal = primal / (1.0 + bn);
    r->relative_dual = dviol / (1.0 + cn);
    r->relative_gap = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
    free(ax);
    free(aty);
}
