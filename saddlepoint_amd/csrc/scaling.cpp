// Host preconditioning: restates crates/saddlepoint-core/src/scaling.rs:53-201
// and the CsrMatrix helpers of problem.rs:88-126, operation for operation, so
// the scaled problem is bitwise the one the reference builds.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>

#include "host.hpp"
#include "saddlepoint_hip.h"

namespace sp {

Csr transpose_of(int64_t nrows, int64_t ncols, const int64_t* row_ptr, const int64_t* col_idx, const double* values) {
    Csr t;
    t.nrows = ncols;
    t.ncols = nrows;
    const int64_t nnz = row_ptr[nrows];
    t.row_ptr.assign((size_t)ncols + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) t.row_ptr[(size_t)col_idx[k] + 1] += 1;
    for (int64_t j = 0; j < ncols; ++j) t.row_ptr[(size_t)j + 1] += t.row_ptr[(size_t)j];
    std::vector<int64_t> next(t.row_ptr.begin(), t.row_ptr.end() - 1);
    t.col_idx.resize((size_t)nnz);
    t.values.resize((size_t)nnz);
    for (int64_t i = 0; i < nrows; ++i)
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
            const int64_t q = next[(size_t)col_idx[k]]++;
            t.col_idx[(size_t)q] = i;
            t.values[(size_t)q] = values[k];
        }
    return t;
}

Csr Csr::transpose() const { return transpose_of(nrows, ncols, row_ptr.data(), col_idx.data(), values.data()); }

void Csr::scale(const std::vector<double>& row, const std::vector<double>& col) {
    for (int64_t i = 0; i < nrows; ++i) {
        const double ri = row[(size_t)i];
        for (int64_t k = row_ptr[(size_t)i]; k < row_ptr[(size_t)i + 1]; ++k)
            values[(size_t)k] *= ri * col[(size_t)col_idx[(size_t)k]];
    }
}

static double inv_sqrt(double v) { return (v > 0.0 && std::isfinite(v)) ? 1.0 / std::sqrt(v) : 1.0; }

double norm2(const double* v, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += v[i] * v[i];
    return std::sqrt(s);
}

double finite_bound_norm(const double* lo, const double* hi, int64_t m) {
    double s = 0.0;
    for (int64_t i = 0; i < m; ++i) {
        const double l = std::isfinite(lo[i]) ? lo[i] : 0.0;
        const double u = std::isfinite(hi[i]) ? hi[i] : 0.0;
        s += (l == u) ? l * l : l * l + u * u;
    }
    return std::sqrt(s);
}

namespace {

// Smallest and largest nonzero magnitude per row and column; 0 for an empty
// line, which inv_sqrt maps to a factor of 1 (scaling.rs:159-178).
void min_max(const Csr& a, std::vector<double>& rmin, std::vector<double>& rmax, std::vector<double>& cmin,
             std::vector<double>& cmax) {
    rmin.assign((size_t)a.nrows, INFINITY);
    rmax.assign((size_t)a.nrows, 0.0);
    cmin.assign((size_t)a.ncols, INFINITY);
    cmax.assign((size_t)a.ncols, 0.0);
    for (int64_t i = 0; i < a.nrows; ++i)
        for (int64_t k = a.row_ptr[(size_t)i]; k < a.row_ptr[(size_t)i + 1]; ++k) {
            const size_t j = (size_t)a.col_idx[(size_t)k];
            const double v = std::fabs(a.values[(size_t)k]);
            if (v < rmin[(size_t)i]) rmin[(size_t)i] = v;
            if (v > rmax[(size_t)i]) rmax[(size_t)i] = v;
            if (v < cmin[j]) cmin[j] = v;
            if (v > cmax[j]) cmax[j] = v;
        }
    for (size_t i = 0; i < rmin.size(); ++i)
        if (rmax[i] == 0.0) rmin[i] = 0.0;
    for (size_t j = 0; j < cmin.size(); ++j)
        if (cmax[j] == 0.0) cmin[j] = 0.0;
}

// Per-row and per-column max of |a| (scaling.rs:144-155 with acc.max(v)).
void abs_max(const Csr& a, std::vector<double>& rows, std::vector<double>& cols) {
    rows.assign((size_t)a.nrows, 0.0);
    cols.assign((size_t)a.ncols, 0.0);
    for (int64_t i = 0; i < a.nrows; ++i)
        for (int64_t k = a.row_ptr[(size_t)i]; k < a.row_ptr[(size_t)i + 1]; ++k) {
            const size_t j = (size_t)a.col_idx[(size_t)k];
            const double v = std::fabs(a.values[(size_t)k]);
            if (v > rows[(size_t)i]) rows[(size_t)i] = v;
            if (v > cols[j]) cols[j] = v;
        }
}

// Per-row and per-column Σ |a|^e in CSR order (scaling.rs:144-155 with acc + v.powf(e)).
void abs_pow_sum(const Csr& a, double e, std::vector<double>& rows, std::vector<double>& cols) {
    rows.assign((size_t)a.nrows, 0.0);
    cols.assign((size_t)a.ncols, 0.0);
    for (int64_t i = 0; i < a.nrows; ++i)
        for (int64_t k = a.row_ptr[(size_t)i]; k < a.row_ptr[(size_t)i + 1]; ++k) {
            const size_t j = (size_t)a.col_idx[(size_t)k];
            const double t = std::pow(std::fabs(a.values[(size_t)k]), e);
            rows[(size_t)i] = rows[(size_t)i] + t;
            cols[j] = cols[j] + t;
        }
}

void apply(Csr& a, std::vector<double>& rs, std::vector<double>& cs, const std::vector<double>& r,
           const std::vector<double>& c) {
    a.scale(r, c);
    for (size_t i = 0; i < rs.size(); ++i) rs[i] *= r[i];
    for (size_t j = 0; j < cs.size(); ++j) cs[j] *= c[j];
}

}  // namespace

ScaledProblem ScaledProblem::build(const LpView& p, const ScalingParams& params) {
    const int64_t m = p.m, n = p.n;
    ScaledProblem s;
    s.a.nrows = m;
    s.a.ncols = n;
    s.a.row_ptr.assign(p.row_ptr, p.row_ptr + m + 1);
    s.a.col_idx.assign(p.col_idx, p.col_idx + p.nnz);
    s.a.values.assign(p.values, p.values + p.nnz);
    s.row_scale.assign((size_t)m, 1.0);
    s.col_scale.assign((size_t)n, 1.0);
    const std::vector<double> ones_m((size_t)m, 1.0), ones_n((size_t)n, 1.0);
    std::vector<double> rmin, rmax, cmin, cmax, r((size_t)m), c((size_t)n);

    for (int64_t pass = 0; pass < params.geometric_mean_passes; ++pass) {
        min_max(s.a, rmin, rmax, cmin, cmax);
        for (int64_t i = 0; i < m; ++i) r[(size_t)i] = inv_sqrt(rmin[(size_t)i] * rmax[(size_t)i]);
        apply(s.a, s.row_scale, s.col_scale, r, ones_n);
        min_max(s.a, rmin, rmax, cmin, cmax);
        for (int64_t j = 0; j < n; ++j) c[(size_t)j] = inv_sqrt(cmin[(size_t)j] * cmax[(size_t)j]);
        apply(s.a, s.row_scale, s.col_scale, ones_m, c);
    }
    for (int64_t it = 0; it < params.ruiz_iterations; ++it) {
        abs_max(s.a, rmax, cmax);
        for (int64_t i = 0; i < m; ++i) r[(size_t)i] = inv_sqrt(rmax[(size_t)i]);
        for (int64_t j = 0; j < n; ++j) c[(size_t)j] = inv_sqrt(cmax[(size_t)j]);
        apply(s.a, s.row_scale, s.col_scale, r, c);
    }
    if (params.has_pock_chambolle) {
        const double alpha = params.pock_chambolle_alpha;
        abs_pow_sum(s.a, alpha, rmax, cmax);
        for (int64_t i = 0; i < m; ++i) r[(size_t)i] = inv_sqrt(rmax[(size_t)i]);
        abs_pow_sum(s.a, 2.0 - alpha, rmin, cmax);
        for (int64_t j = 0; j < n; ++j) c[(size_t)j] = inv_sqrt(cmax[(size_t)j]);
        apply(s.a, s.row_scale, s.col_scale, r, c);
    }

    finish_scaling(p, params, s);
    s.at = s.a.transpose();
    return s;
}

void finish_scaling(const LpView& p, const ScalingParams& params, ScaledProblem& s, const int64_t* rank) {
    const int64_t m = p.m, n = p.n;
    s.c.resize((size_t)n);
    s.lv.resize((size_t)n);
    s.uv.resize((size_t)n);
    s.lc.resize((size_t)m);
    s.uc.resize((size_t)m);
    for (int64_t j = 0; j < n; ++j) {
        s.c[(size_t)j] = p.c[j] * s.col_scale[(size_t)j];
        s.lv[(size_t)j] = p.lv[j] / s.col_scale[(size_t)j];
        s.uv[(size_t)j] = p.uv[j] / s.col_scale[(size_t)j];
    }
    for (int64_t i = 0; i < m; ++i) {
        s.lc[(size_t)i] = p.lc[i] * s.row_scale[(size_t)i];
        s.uc[(size_t)i] = p.uc[i] * s.row_scale[(size_t)i];
    }
    if (params.bound_objective_rescaling) {
        s.bound_scale = 1.0 / (1.0 + finite_bound_norm(s.lc.data(), s.uc.data(), m));
        double cn;
        if (rank) {
            double acc = 0.0;
            for (int64_t j = 0; j < n; ++j) {
                const double v = s.c[(size_t)rank[j]];
                acc += v * v;
            }
            cn = std::sqrt(acc);
        } else {
            cn = norm2(s.c.data(), n);
        }
        s.obj_scale = 1.0 / (1.0 + cn);
        for (auto* v : {&s.lc, &s.uc, &s.lv, &s.uv})
            for (double& e : *v) e *= s.bound_scale;
        for (double& e : s.c) e *= s.obj_scale;
    }
}

std::vector<double> ScaledProblem::unscale_primal(const std::vector<double>& xs) const {
    std::vector<double> x(xs.size());
    for (size_t j = 0; j < xs.size(); ++j) x[j] = xs[j] * col_scale[j] / bound_scale;
    return x;
}

std::vector<double> ScaledProblem::unscale_dual(const std::vector<double>& ys) const {
    std::vector<double> y(ys.size());
    for (size_t i = 0; i < ys.size(); ++i) y[i] = ys[i] * row_scale[i] / obj_scale;
    return y;
}

std::vector<double> ScaledProblem::primal_residual_factors() const {
    std::vector<double> f(row_scale.size());
    for (size_t i = 0; i < f.size(); ++i) f[i] = 1.0 / (row_scale[i] * bound_scale);
    return f;
}

std::vector<double> ScaledProblem::dual_residual_factors() const {
    std::vector<double> f(col_scale.size());
    for (size_t j = 0; j < f.size(); ++j) f[j] = 1.0 / (col_scale[j] * obj_scale);
    return f;
}

ColumnOrder popularity_order(const LpView& p) {
    std::vector<int64_t> count((size_t)p.n, 0);
    for (int64_t k = 0; k < p.nnz; ++k) count[(size_t)p.col_idx[k]] += 1;
    return popularity_order_of_counts(count);
}

RowOrder long_rows_first(const LpView& p, const ColumnOrder& cols, int64_t long_row) {
    RowOrder o;
    const int64_t m = p.m, n = p.n;
    o.order.reserve((size_t)m);
    // The short rows by key (smallest relabeled column; n for an empty row),
    // a counting sort that keeps the index order among equal keys.
    std::vector<int64_t> key((size_t)m, -1), start((size_t)n + 2, 0);
    for (int64_t i = 0; i < m; ++i) {
        const int64_t k0 = p.row_ptr[i], k1 = p.row_ptr[i + 1];
        if (k1 - k0 > long_row) {
            o.order.push_back(i);
            continue;
        }
        int64_t kmin = n;
        for (int64_t k = k0; k < k1; ++k) {
            const int64_t j = cols.empty() ? p.col_idx[k] : cols.rank[(size_t)p.col_idx[k]];
            kmin = std::min(kmin, j);
        }
        key[(size_t)i] = kmin;
        start[(size_t)kmin + 1] += 1;
    }
    // The long rows by decreasing length, ties by index.
    std::stable_sort(o.order.begin(), o.order.end(), [&](int64_t a, int64_t b) {
        return p.row_ptr[a + 1] - p.row_ptr[a] > p.row_ptr[b + 1] - p.row_ptr[b];
    });
    for (size_t b = 1; b < start.size(); ++b) start[b] += start[b - 1];
    const size_t base = o.order.size();
    o.order.resize((size_t)m);
    for (int64_t i = 0; i < m; ++i)
        if (key[(size_t)i] >= 0) o.order[base + (size_t)start[(size_t)key[(size_t)i]]++] = i;
    return o;
}

ColumnOrder popularity_order_of_counts(const std::vector<int64_t>& count) {
    ColumnOrder o;
    const int64_t n = (int64_t)count.size();
    // A counting sort on the counts, largest first; columns of equal count
    // keep their index order.
    int64_t top = 0;
    for (int64_t c : count) top = std::max(top, c);
    std::vector<int64_t> start((size_t)top + 2, 0);
    for (int64_t c : count) start[(size_t)(top - c) + 1] += 1;
    for (size_t b = 1; b < start.size(); ++b) start[b] += start[b - 1];
    o.order.resize((size_t)n);
    o.rank.resize((size_t)n);
    for (int64_t j = 0; j < n; ++j) {
        const int64_t q = start[(size_t)(top - count[(size_t)j])]++;
        o.order[(size_t)q] = j;
        o.rank[(size_t)j] = q;
    }
    return o;
}

RelabeledLp relabel_vectors(const LpView& p, const ColumnOrder& o) {
    RelabeledLp r;
    r.v = p;
    r.c.resize((size_t)p.n);
    r.lv.resize((size_t)p.n);
    r.uv.resize((size_t)p.n);
    for (int64_t q = 0; q < p.n; ++q) {
        const int64_t j = o.order[(size_t)q];
        r.c[(size_t)q] = p.c[j];
        r.lv[(size_t)q] = p.lv[j];
        r.uv[(size_t)q] = p.uv[j];
    }
    r.v.c = r.c.data();
    r.v.lv = r.lv.data();
    r.v.uv = r.uv.data();
    return r;
}

RelabeledLp relabel(const LpView& p, const ColumnOrder& o) {
    RelabeledLp r;
    r.v = p;
    r.col_idx.resize((size_t)p.nnz);
    for (int64_t k = 0; k < p.nnz; ++k) r.col_idx[(size_t)k] = o.rank[(size_t)p.col_idx[k]];
    r.c.resize((size_t)p.n);
    r.lv.resize((size_t)p.n);
    r.uv.resize((size_t)p.n);
    for (int64_t q = 0; q < p.n; ++q) {
        const int64_t j = o.order[(size_t)q];
        r.c[(size_t)q] = p.c[j];
        r.lv[(size_t)q] = p.lv[j];
        r.uv[(size_t)q] = p.uv[j];
    }
    r.v.col_idx = r.col_idx.data();
    r.v.c = r.c.data();
    r.v.lv = r.lv.data();
    r.v.uv = r.uv.data();
    return r;
}

void relabel(ScaledProblem& s, const ColumnOrder& o) {
    const int64_t n = s.a.ncols;
    for (auto& j : s.a.col_idx) j = o.rank[(size_t)j];
    // A'ᵀ: row q of the relabeled transpose is row order[q] of the old one.
    Csr t;
    t.nrows = s.at.nrows;
    t.ncols = s.at.ncols;
    t.row_ptr.resize((size_t)n + 1);
    t.col_idx.resize(s.at.col_idx.size());
    t.values.resize(s.at.values.size());
    t.row_ptr[0] = 0;
    for (int64_t q = 0; q < n; ++q) {
        const int64_t j = o.order[(size_t)q];
        const int64_t b = s.at.row_ptr[(size_t)j], e = s.at.row_ptr[(size_t)j + 1];
        const int64_t at = t.row_ptr[(size_t)q];
        std::copy(s.at.col_idx.begin() + b, s.at.col_idx.begin() + e, t.col_idx.begin() + at);
        std::copy(s.at.values.begin() + b, s.at.values.begin() + e, t.values.begin() + at);
        t.row_ptr[(size_t)q + 1] = at + (e - b);
    }
    s.at = std::move(t);
    auto permute = [&](std::vector<double>& v) {
        std::vector<double> w(v.size());
        for (int64_t q = 0; q < n; ++q) w[(size_t)q] = v[(size_t)o.order[(size_t)q]];
        v.swap(w);
    };
    permute(s.c);
    permute(s.lv);
    permute(s.uv);
    permute(s.col_scale);
}

void validate(const LpView& p) {
    char buf[160];
    if (p.m < 0 || p.n < 0 || p.nnz < 0 || !p.row_ptr || (p.nnz > 0 && (!p.col_idx || !p.values)))
        throw Error(SP_ERR_PROBLEM, "A has an invalid shape");
    if (p.row_ptr[0] != 0 || p.row_ptr[p.m] != p.nnz) throw Error(SP_ERR_PROBLEM, "A row pointers do not span nnz");
    for (int64_t i = 0; i < p.m; ++i)
        if (p.row_ptr[i + 1] < p.row_ptr[i]) throw Error(SP_ERR_PROBLEM, "A row pointers decrease");
    for (int64_t k = 0; k < p.nnz; ++k)
        if (p.col_idx[k] < 0 || p.col_idx[k] >= p.n) throw Error(SP_ERR_PROBLEM, "A column index out of range");
    struct Pair {
        const char* what;
        const double *lo, *hi;
        int64_t len;
    };
    for (const Pair& q : {Pair{"variable", p.lv, p.uv, p.n}, Pair{"constraint", p.lc, p.uc, p.m}}) {
        for (int64_t i = 0; i < q.len; ++i) {
            if (std::isnan(q.lo[i]) || std::isnan(q.hi[i])) {
                std::snprintf(buf, sizeof buf, "invalid problem: %s %lld is NaN", q.what, (long long)i);
                throw Error(SP_ERR_PROBLEM, buf);
            }
            if (q.lo[i] > q.hi[i]) {
                std::snprintf(buf, sizeof buf, "invalid problem: %s %lld has lower bound %g above upper bound %g",
                              q.what, (long long)i, q.lo[i], q.hi[i]);
                throw Error(SP_ERR_PROBLEM, buf);
            }
        }
    }
    for (int64_t j = 0; j < p.n; ++j)
        if (std::isnan(p.c[j])) {
            std::snprintf(buf, sizeof buf, "invalid problem: coefficient %lld is NaN", (long long)j);
            throw Error(SP_ERR_PROBLEM, buf);
        }
    for (int64_t k = 0; k < p.nnz; ++k)
        if (std::isnan(p.values[k])) {
            std::snprintf(buf, sizeof buf, "invalid problem: coefficient %lld is NaN", (long long)(p.n + k));
            throw Error(SP_ERR_PROBLEM, buf);
        }
}

}  // namespace sp
