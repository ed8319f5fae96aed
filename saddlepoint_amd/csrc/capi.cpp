// The C ABI of include/saddlepoint_hip.h: argument checks, error codes and
// messages around the HIP backend (backend.cpp) and the driver (solver.hpp).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>

#include "backend.hpp"
#include "saddlepoint_hip.h"
#include "driver.hpp"
#include "solver.hpp"

namespace {

thread_local std::string g_create_error;

// Runs f, mapping exceptions to an error code and a message on the backend.
template <class F>
int guarded(sp_backend* b, F&& f) {
    if (!b) return SP_ERR_INVALID;
    try {
        f();
        sp::set_error(b, "");
        return SP_OK;
    } catch (const sp::Error& e) {
        sp::set_error(b, e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        sp::set_error(b, "host out of memory");
        return SP_ERR_MEMORY;
    } catch (const std::exception& e) {
        sp::set_error(b, e.what());
        return SP_ERR_INVALID;
    }
}

void check_problem(const sp_backend* b, const sp_problem* p) {
    if (!p) throw sp::Error(SP_ERR_INVALID, "null problem");
    if (p->b != b) throw sp::Error(SP_ERR_INVALID, "problem belongs to another backend");
}

sp::LpView view_of(const sp_lp_view* lp) {
    if (!lp) throw sp::Error(SP_ERR_INVALID, "null problem");
    sp::LpView v;
    v.m = lp->a.nrows;
    v.n = lp->a.ncols;
    v.nnz = lp->a.nnz;
    v.row_ptr = lp->a.row_ptr;
    v.col_idx = lp->a.col_idx;
    v.values = lp->a.values;
    v.c = lp->c;
    v.lc = lp->lc;
    v.uc = lp->uc;
    v.lv = lp->lv;
    v.uv = lp->uv;
    v.c0 = lp->c0;
    v.maximize = lp->maximize != 0;
    if (v.m < 0 || v.n < 0 || (v.n && (!v.c || !v.lv || !v.uv)) || (v.m && (!v.lc || !v.uc)))
        throw sp::Error(SP_ERR_INVALID, "null array in problem");
    return v;
}

sp::ScalingParams scaling_of(const sp_params& p) { return sp::scaling_params_of(p); }

}  // namespace

using sp::DriverBackend;

struct sp_solver {
    sp_backend* b;
    DriverBackend db;
    sp::Solver<DriverBackend> solver;
    double setup_s;
    sp_solver(sp_backend* bb, const sp::LpView& v, const sp_params& p, sp::ScaledProblem s, double setup)
        : b(bb), db{{bb}}, solver(db, v, p, std::move(s)), setup_s(setup) {}
};

extern "C" {

int sp_backend_create(int ordinal, sp_backend** out) {
    if (!out) return SP_ERR_INVALID;
    *out = nullptr;
    try {
        *out = sp_backend_open(ordinal);
        g_create_error.clear();
        return SP_OK;
    } catch (const sp::Error& e) {
        g_create_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return SP_ERR_UNAVAILABLE;
    }
}

void sp_backend_destroy(sp_backend* b) { sp_backend_close(b); }

const char* sp_backend_name(const sp_backend* b) { return b ? b->name.c_str() : ""; }

const char* sp_last_error(const sp_backend* b) { return b ? sp::last_error(b) : g_create_error.c_str(); }

int sp_problem_load(sp_backend* b, const sp_scaled_view* p, sp_problem** out) {
    return guarded(b, [&] {
        if (!p || !out) throw sp::Error(SP_ERR_INVALID, "load: null argument");
        *out = nullptr;
        *out = sp::Hip{b}.load_view(*p);
    });
}

void sp_problem_free(sp_problem* p) { sp::free_problem(p); }

int sp_vec_upload(sp_backend* b, const double* data, int64_t len, sp_vec** out) {
    return guarded(b, [&] {
        if (!out || (len && !data)) throw sp::Error(SP_ERR_INVALID, "upload: null argument");
        *out = sp::Hip{b}.upload(data, len);
    });
}

int sp_vec_zeros(sp_backend* b, int64_t len, sp_vec** out) {
    return guarded(b, [&] {
        if (!out) throw sp::Error(SP_ERR_INVALID, "zeros: null argument");
        *out = sp::Hip{b}.zeros(len);
    });
}

int sp_vec_download(sp_backend* b, const sp_vec* v, double* out, int64_t len) {
    return guarded(b, [&] {
        if (len && !out) throw sp::Error(SP_ERR_INVALID, "download: null argument");
        sp::Hip{b}.download(v, out, len);
    });
}

int sp_vec_copy(sp_backend* b, const sp_vec* src, sp_vec* dst) {
    return guarded(b, [&] {
        if (!src) throw sp::Error(SP_ERR_INVALID, "copy: null vector");
        sp::Hip{b}.copy(src, dst);
    });
}

int64_t sp_vec_len(const sp_vec* v) { return v ? v->len : -1; }

void sp_vec_free(sp_vec* v) { sp::free_vec(v); }

int sp_spmv(sp_backend* b, const sp_problem* p, const sp_vec* x, sp_vec* out) {
    return guarded(b, [&] {
        check_problem(b, p);
        sp::Hip{b}.spmv(p, x, out);
    });
}

int sp_spmv_t(sp_backend* b, const sp_problem* p, const sp_vec* y, sp_vec* out) {
    return guarded(b, [&] {
        check_problem(b, p);
        sp::Hip{b}.spmv_t(p, y, out);
    });
}

int sp_primal_step(sp_backend* b, const sp_problem* p, const sp_vec* x, const sp_vec* aty, double tau, sp_vec* x_t,
                   sp_vec* x_bar) {
    return guarded(b, [&] {
        check_problem(b, p);
        sp::Hip{b}.primal_step(p, x, aty, tau, x_t, x_bar);
    });
}

int sp_dual_step(sp_backend* b, const sp_problem* p, const sp_vec* y, const sp_vec* ax_bar, double sigma,
                 sp_vec* y_t) {
    return guarded(b, [&] {
        check_problem(b, p);
        sp::Hip{b}.dual_step(p, y, ax_bar, sigma, y_t);
    });
}

int sp_halpern_step(sp_backend* b, sp_vec* z, const sp_vec* z_t, const sp_vec* anchor, double w, double rho) {
    return guarded(b, [&] { sp::Hip{b}.halpern_step(z, z_t, anchor, w, rho); });
}

int sp_distance_sq(sp_backend* b, const sp_vec* a, const sp_vec* bv, double* out) {
    return guarded(b, [&] {
        if (!out) throw sp::Error(SP_ERR_INVALID, "null output");
        *out = sp::Hip{b}.distance_sq(a, bv);
    });
}

int sp_sub(sp_backend* b, const sp_vec* a, const sp_vec* bv, sp_vec* out) {
    return guarded(b, [&] { sp::Hip{b}.sub(a, bv, out); });
}

int sp_unscale_primal(sp_backend* b, const sp_vec* xs, const sp_vec* col_scale, double bound_scale, sp_vec* out) {
    return guarded(b, [&] {
        if (!xs || !col_scale) throw sp::Error(SP_ERR_INVALID, "unscale_primal: null vector");
        sp::Hip{b}.unscale(xs, col_scale->d, col_scale->len, bound_scale, nullptr, false, out);
    });
}

int sp_unscale_dual(sp_backend* b, const sp_vec* ys, const sp_vec* row_scale, double obj_scale, sp_vec* out) {
    return guarded(b, [&] {
        if (!ys || !row_scale) throw sp::Error(SP_ERR_INVALID, "unscale_dual: null vector");
        sp::Hip{b}.unscale(ys, row_scale->d, row_scale->len, obj_scale, nullptr, true, out);
    });
}

int sp_norm_sq(sp_backend* b, const sp_vec* v, double* out) {
    return guarded(b, [&] {
        if (!out) throw sp::Error(SP_ERR_INVALID, "null output");
        *out = sp::Hip{b}.norm_sq(v);
    });
}

int sp_dot(sp_backend* b, const sp_vec* a, const sp_vec* bv, double* out) {
    return guarded(b, [&] {
        if (!out) throw sp::Error(SP_ERR_INVALID, "null output");
        *out = sp::Hip{b}.dot(a, bv);
    });
}

int sp_scale(sp_backend* b, sp_vec* v, double alpha) {
    return guarded(b, [&] { sp::Hip{b}.scale(v, alpha); });
}

int sp_primal_stats(sp_backend* b, const sp_problem* p, const sp_vec* x, const sp_vec* ax, sp_primal_stats_t* out) {
    return guarded(b, [&] {
        check_problem(b, p);
        if (!out) throw sp::Error(SP_ERR_INVALID, "null output");
        *out = sp::Hip{b}.primal_stats(p, x, ax);
    });
}

int sp_dual_stats(sp_backend* b, const sp_problem* p, const sp_vec* y, const sp_vec* aty, sp_dual_stats_t* out) {
    return guarded(b, [&] {
        check_problem(b, p);
        if (!out) throw sp::Error(SP_ERR_INVALID, "null output");
        *out = sp::Hip{b}.dual_stats(p, y, aty);
    });
}

int sp_step_dual_residual_sq(sp_backend* b, const sp_problem* p, const sp_vec* x, const sp_vec* x_t,
                             const sp_vec* aty, const sp_vec* aty_t, double tau, double* out) {
    return guarded(b, [&] {
        check_problem(b, p);
        if (!out) throw sp::Error(SP_ERR_INVALID, "null output");
        *out = sp::Hip{b}.step_dual_residual_sq(p, x, x_t, aty, aty_t, tau);
    });
}

int sp_primal_ray_stats(sp_backend* b, const sp_problem* p, const sp_vec* dx, const sp_vec* adx, sp_ray_stats_t* out) {
    return guarded(b, [&] {
        check_problem(b, p);
        if (!out) throw sp::Error(SP_ERR_INVALID, "null output");
        *out = sp::Hip{b}.primal_ray_stats(p, dx, adx);
    });
}

int sp_dual_ray_stats(sp_backend* b, const sp_problem* p, const sp_vec* dy, const sp_vec* atdy, sp_ray_stats_t* out) {
    return guarded(b, [&] {
        check_problem(b, p);
        if (!out) throw sp::Error(SP_ERR_INVALID, "null output");
        *out = sp::Hip{b}.dual_ray_stats(p, dy, atdy);
    });
}

void sp_default_params(sp_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    // SolverParams::default (solver.rs:70-93) and ScalingParams::default (scaling.rs:25-34).
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
    p->power_iterations = 200;
    p->geometric_mean_passes = 12;
    p->ruiz_iterations = 10;
    p->has_pock_chambolle = 1;
    p->pock_chambolle_alpha = 1.0;
    p->bound_objective_rescaling = 1;
    p->termination = 0;
    p->fused = 1;
    p->log_every_checks = 0;
    p->precision = 0;
    p->spmv_autotune = 0;
    p->gpu_scaling = 1;
    p->reorder_columns = 1;
    p->column_panels = -1;
    p->hot_table = -1;
    p->reorder_rows = -1;
}

int sp_solve(sp_backend* b, const sp_lp_view* lp, const sp_params* params, sp_solution* out, double* x_out,
             double* y_out) {
    return guarded(b, [&] {
        if (!params || !out) throw sp::Error(SP_ERR_INVALID, "solve: null argument");
        if (params->check_every == 0) throw sp::Error(SP_ERR_INVALID, "solve: check_every must be positive");
        const auto t0 = std::chrono::steady_clock::now();
        const sp::LpView v = view_of(lp);
        sp::validate(v);
        sp::ScaledProblem scaled =
            params->gpu_scaling ? sp::ScaledProblem{} : sp::ScaledProblem::build(v, scaling_of(*params));
        DriverBackend db{{b}};
        sp::Solver<DriverBackend> solver(db, v, *params, std::move(scaled));
        const double setup = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = solver.run(x_out, y_out);
        out->setup_time_s = setup;
        out->total_time_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
}

int sp_solver_create(sp_backend* b, const sp_lp_view* lp, const sp_params* params, sp_solver** out) {
    return guarded(b, [&] {
        if (!params || !out) throw sp::Error(SP_ERR_INVALID, "solver_create: null argument");
        if (params->check_every == 0) throw sp::Error(SP_ERR_INVALID, "solver_create: check_every must be positive");
        *out = nullptr;
        const auto t0 = std::chrono::steady_clock::now();
        const sp::LpView v = view_of(lp);
        sp::validate(v);
        sp::ScaledProblem scaled =
            params->gpu_scaling ? sp::ScaledProblem{} : sp::ScaledProblem::build(v, scaling_of(*params));
        auto* s = new sp_solver(b, v, *params, std::move(scaled),
                                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        try {
            s->solver.start();
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int sp_solver_advance(sp_solver* s, uint64_t iterations, int* done) {
    if (!s) return SP_ERR_INVALID;
    return guarded(s->b, [&] {
        const bool d = s->solver.advance(iterations);
        if (done) *done = d ? 1 : 0;
    });
}

int sp_solver_result(sp_solver* s, sp_solution* out, double* x_out, double* y_out) {
    if (!s) return SP_ERR_INVALID;
    return guarded(s->b, [&] {
        if (!out) throw sp::Error(SP_ERR_INVALID, "solver_result: null output");
        *out = s->solver.finish(x_out, y_out);
        out->setup_time_s = s->setup_s;
        out->total_time_s = s->setup_s + out->solve_time_s;
    });
}

void sp_solver_free(sp_solver* s) { delete s; }

int sp_backend_profile(sp_backend* b, int enable) {
    return guarded(b, [&] {
        sp::profile_collect(b);
        b->profiling = enable != 0;
        if (b->profiling) {
            std::fill(b->kernel_ms.begin(), b->kernel_ms.end(), 0.0);
            std::fill(b->kernel_launches.begin(), b->kernel_launches.end(), 0);
        }
    });
}

int sp_backend_profile_read(sp_backend* b, int index, char* name, int name_len, double* total_ms,
                            int64_t* launches) {
    return guarded(b, [&] {
        sp::profile_collect(b);
        if (index < 0 || index >= (int)b->kernel_names.size()) throw sp::Error(SP_ERR_INVALID, "no such kernel");
        if (name && name_len > 0) {
            std::strncpy(name, b->kernel_names[(size_t)index].c_str(), (size_t)name_len - 1);
            name[name_len - 1] = 0;
        }
        if (total_ms) *total_ms = b->kernel_ms[(size_t)index];
        if (launches) *launches = b->kernel_launches[(size_t)index];
    });
}

int sp_probe(sp_backend* b, int kind, int64_t n, int64_t table_len, int reps, double* ms_per_launch) {
    return guarded(b, [&] {
        if (!ms_per_launch || kind < 0 || kind > 8) throw sp::Error(SP_ERR_INVALID, "probe: bad argument");
        if (table_len > INT32_MAX) throw sp::Error(SP_ERR_TOO_LARGE, "probe: table beyond int32 indices");
        *ms_per_launch = sp::Hip{b}.probe(kind, n, table_len, reps, 1);
    });
}

int sp_gather_replay_time(sp_backend* b, const sp_problem* p, int transpose, int fp32, int reps, double* ms) {
    return guarded(b, [&] {
        check_problem(b, p);
        if (!ms) throw sp::Error(SP_ERR_INVALID, "gather_replay_time: null output");
        *ms = sp::Hip{b}.gather_replay_time(p, transpose != 0, fp32 != 0, reps);
    });
}

const sp_problem* sp_solver_problem(const sp_solver* s) { return s ? s->solver.problem() : nullptr; }

int sp_solver_kernel_time(sp_solver* s, int side, int reps, double* ms) {
    if (!s) return SP_ERR_INVALID;
    return guarded(s->b, [&] {
        if (!ms) throw sp::Error(SP_ERR_INVALID, "solver_kernel_time: null output");
        *ms = s->solver.plain_kernel_time(side, reps);
    });
}

int sp_backend_sync(sp_backend* b) {
    return guarded(b, [&] { sp::Hip{b}.sync(); });
}

namespace {
sp_spmv_tuning to_c(const sp::SpmvTuning& t) { return sp_spmv_tuning{t.block_nnz, t.long_row, t.max_rows}; }
sp::SpmvTuning from_c(const sp_spmv_tuning& t) {
    if (t.block_nnz < 1 || t.block_nnz > 2048 || t.long_row < 1 || t.long_row > t.block_nnz || t.max_rows < 1 ||
        t.max_rows > 1024)
        throw sp::Error(SP_ERR_INVALID, "spmv tuning out of range");
    sp::SpmvTuning r;
    r.block_nnz = t.block_nnz;
    r.long_row = t.long_row;
    r.max_rows = t.max_rows;
    return r;
}
}  // namespace

void sp_spmv_tuning_default(sp_spmv_tuning* out) {
    if (out) *out = to_c(sp::SpmvTuning{});
}

void sp_spmv_tuning_for(const int64_t* row_ptr, int64_t nrows, sp_spmv_tuning* out) {
    if (!out) return;
    *out = (row_ptr && nrows >= 0) ? to_c(sp::autotuned_plan(row_ptr, nrows)) : to_c(sp::SpmvTuning{});
}

int sp_problem_load_tuned(sp_backend* b, const sp_scaled_view* p, const sp_spmv_tuning* a, const sp_spmv_tuning* at,
                          sp_problem** out) {
    return guarded(b, [&] {
        if (!p || !out) throw sp::Error(SP_ERR_INVALID, "load: null argument");
        *out = nullptr;
        const sp::SpmvTuning ta = a ? from_c(*a) : sp::SpmvTuning{};
        const sp::SpmvTuning tat = at ? from_c(*at) : sp::SpmvTuning{};
        *out = sp::Hip{b}.load_view(*p, &ta, &tat);
    });
}

int sp_spmv_time(sp_backend* b, const sp_problem* p, int transpose, int reps, double* ms) {
    return guarded(b, [&] {
        if (!p || !ms) throw sp::Error(SP_ERR_INVALID, "spmv_time: null argument");
        *ms = sp::Hip{b}.spmv_time(p, transpose != 0, reps);
    });
}

int sp_problem_set_panels(sp_backend* b, sp_problem* p, int shift) {
    return guarded(b, [&] {
        check_problem(b, p);
        sp::Hip{b}.set_panels(p, shift);
    });
}

int sp_scale_problem(const sp_lp_view* lp, const sp_params* params, double* a_values, double* c, double* lc,
                     double* uc, double* lv, double* uv, double* row_scale, double* col_scale, double* bound_scale,
                     double* obj_scale) {
    try {
        if (!params) return SP_ERR_INVALID;
        const sp::LpView v = view_of(lp);
        sp::validate(v);
        const sp::ScaledProblem s = sp::ScaledProblem::build(v, scaling_of(*params));
        std::copy(s.a.values.begin(), s.a.values.end(), a_values);
        std::copy(s.c.begin(), s.c.end(), c);
        std::copy(s.lc.begin(), s.lc.end(), lc);
        std::copy(s.uc.begin(), s.uc.end(), uc);
        std::copy(s.lv.begin(), s.lv.end(), lv);
        std::copy(s.uv.begin(), s.uv.end(), uv);
        std::copy(s.row_scale.begin(), s.row_scale.end(), row_scale);
        std::copy(s.col_scale.begin(), s.col_scale.end(), col_scale);
        *bound_scale = s.bound_scale;
        *obj_scale = s.obj_scale;
        return SP_OK;
    } catch (const sp::Error& e) {
        g_create_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return SP_ERR_INVALID;
    }
}

int sp_lp_read_mps(const char* path, sp_lp** out) {
    try {
        if (!path || !out) throw sp::Error(SP_ERR_INVALID, "read_mps: null argument");
        *out = nullptr;
        std::FILE* fp = std::fopen(path, "rb");
        if (!fp) throw sp::Error(SP_ERR_INVALID, std::string(path) + ": cannot open");
        std::string text;
        char buf[1 << 16];
        size_t got;
        while ((got = std::fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, got);
        const bool bad = std::ferror(fp) != 0;
        std::fclose(fp);
        if (bad) throw sp::Error(SP_ERR_INVALID, std::string(path) + ": read error");
        *out = sp::read_mps_text(text.data(), text.size());
        if ((*out)->name.empty()) {
            std::string stem(path);
            const size_t slash = stem.find_last_of('/');
            if (slash != std::string::npos) stem = stem.substr(slash + 1);
            const size_t dot = stem.find('.');
            if (dot != std::string::npos) stem = stem.substr(0, dot);
            (*out)->name = stem;
        }
        g_create_error.clear();
        return SP_OK;
    } catch (const sp::Error& e) {
        g_create_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_create_error = "host out of memory";
        return SP_ERR_MEMORY;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return SP_ERR_INVALID;
    }
}

int sp_lp_view_of(const sp_lp* lp, sp_lp_view* out) {
    if (!lp || !out) return SP_ERR_INVALID;
    out->a = sp_csr_view{lp->m, lp->n, (int64_t)lp->values.size(), lp->row_ptr.data(), lp->col_idx.data(),
                         lp->values.data()};
    out->c = lp->c.data();
    out->lc = lp->lc.data();
    out->uc = lp->uc.data();
    out->lv = lp->lv.data();
    out->uv = lp->uv.data();
    out->c0 = lp->c0;
    out->maximize = lp->maximize ? 1 : 0;
    return SP_OK;
}

const char* sp_lp_name(const sp_lp* lp) { return lp ? lp->name.c_str() : ""; }

void sp_lp_free(sp_lp* lp) { delete lp; }

int sp_transpose_gpu(sp_backend* b, int64_t nrows, int64_t ncols, const int64_t* row_ptr, const int64_t* col_idx,
                     const double* values, int64_t* t_row_ptr, int64_t* t_col_idx, double* t_values) {
    return guarded(b, [&] {
        if (nrows < 0 || ncols < 0 || !row_ptr || !t_row_ptr) throw sp::Error(SP_ERR_INVALID, "transpose: bad argument");
        const int64_t nnz = row_ptr[nrows];
        if (nnz > 0 && (!col_idx || !values || !t_col_idx || !t_values))
            throw sp::Error(SP_ERR_INVALID, "transpose: null array");
        for (int64_t k = 0; k < nnz; ++k)
            if (col_idx[k] < 0 || col_idx[k] >= ncols) throw sp::Error(SP_ERR_INVALID, "transpose: column out of range");
        sp::DevCsr A, T;
        sp::Hip h{b};
        std::vector<int64_t> trp;
        try {
            h.upload(A, row_ptr, col_idx, values, nrows, ncols, nnz, sp::SpmvTuning{});
            h.transpose(A, T, false, trp);
            std::copy(trp.begin(), trp.end(), t_row_ptr);
            if (nnz) {
                std::vector<int> ci((size_t)nnz);
                sp::hip_check(hipMemcpy(ci.data(), T.ci, 4 * (size_t)nnz, hipMemcpyDeviceToHost), "D2H");
                sp::hip_check(hipMemcpy(t_values, T.va, 8 * (size_t)nnz, hipMemcpyDeviceToHost), "D2H");
                for (int64_t k = 0; k < nnz; ++k) t_col_idx[k] = ci[(size_t)k];
            }
        } catch (...) {
            sp::free_csr_public(A);
            sp::free_csr_public(T);
            throw;
        }
        sp::free_csr_public(A);
        sp::free_csr_public(T);
    });
}

int sp_check(sp_backend* b, const sp_lp_view* lp, const double* x, const double* y, sp_kkt_report* out) {
    return guarded(b, [&] {
        if (!out) throw sp::Error(SP_ERR_INVALID, "check: null output");
        const sp::LpView v = view_of(lp);
        sp::validate(v);
        *out = sp::Hip{b}.check(v, x, y);
    });
}

int sp_scale_problem_gpu(sp_backend* b, const sp_lp_view* lp, const sp_params* params, double* a_values,
                         double* at_values, double* c, double* lc, double* uc, double* lv, double* uv,
                         double* row_scale, double* col_scale, double* bound_scale, double* obj_scale) {
    return guarded(b, [&] {
        if (!params || !a_values || !at_values || !c || !lc || !uc || !lv || !uv || !row_scale || !col_scale ||
            !bound_scale || !obj_scale)
            throw sp::Error(SP_ERR_INVALID, "scale_problem_gpu: null argument");
        const sp::LpView v = view_of(lp);
        sp::validate(v);
        sp::ScaledProblem s;
        std::unique_ptr<sp_problem, void (*)(sp_problem*)> p(
            sp::Hip{b}.load_scaled(v, scaling_of(*params), params->spmv_autotune != 0, s), sp::free_problem);
        if (v.nnz) {
            sp::hip_check(hipMemcpy(a_values, p->a.va, 8 * (size_t)v.nnz, hipMemcpyDeviceToHost), "D2H");
            sp::hip_check(hipMemcpy(at_values, p->at.va, 8 * (size_t)v.nnz, hipMemcpyDeviceToHost), "D2H");
        }
        std::copy(s.c.begin(), s.c.end(), c);
        std::copy(s.lc.begin(), s.lc.end(), lc);
        std::copy(s.uc.begin(), s.uc.end(), uc);
        std::copy(s.lv.begin(), s.lv.end(), lv);
        std::copy(s.uv.begin(), s.uv.end(), uv);
        std::copy(s.row_scale.begin(), s.row_scale.end(), row_scale);
        std::copy(s.col_scale.begin(), s.col_scale.end(), col_scale);
        *bound_scale = s.bound_scale;
        *obj_scale = s.obj_scale;
    });
}

}  // extern "C"
