// HIP backend implementation: device memory, the gfx950 code object, launch
// shapes and the Backend methods (crates/saddlepoint-core/src/backend.rs:43-137,
// behaviour of crates/saddlepoint-cpu/src/lib.rs:101-260).
#include "backend.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstddef>
#include <memory>
#include <climits>
#include <unordered_map>
#include <cstring>

namespace sp {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
            throw Error(SP_ERR_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
        throw Error(SP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}

namespace {
// Per thread: backend -> (its serial, message). The serial tells a backend
// apart from an earlier one that another thread's entry still names at the
// same address.
thread_local std::unordered_map<const sp_backend*, std::pair<uint64_t, std::string>> t_errors;
std::atomic<uint64_t> g_backend_serial{0};
}  // namespace

uint64_t next_backend_serial() { return ++g_backend_serial; }

void set_error(const sp_backend* b, const std::string& msg) {
    if (msg.empty()) {
        t_errors.erase(b);
        return;
    }
    t_errors[b] = {b->serial, msg};
}

const char* last_error(const sp_backend* b) {
    auto it = t_errors.find(b);
    return it == t_errors.end() || it->second.first != b->serial ? "" : it->second.second.c_str();
}

void forget_errors(const sp_backend* b) { t_errors.erase(b); }

namespace {

constexpr int kWG = 256;
constexpr int kChunk = 4096;  // SP_CHUNK in kernels.hip
constexpr int kStreamNnzF64 = 2048;  // SP_NNZ_F64 in kernels.hip
constexpr int kStreamRows = 4 * 256;  // SP_RPT * WG in kernels.hip

template <class T>
T* dmalloc(size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    hip_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    return static_cast<T*>(p);
}

template <class T>
T* dupload(hipStream_t s, const T* src, size_t count) {
    T* d = dmalloc<T>(count);
    if (count) hip_check(hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    return d;
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

// Launches a kernel of the code object; every argument is passed by address,
// as hipModuleLaunchKernel expects, so argument types must match the kernel's.
hipEvent_t take_event(sp_backend* b) {
    if (!b->event_pool.empty()) {
        hipEvent_t e = b->event_pool.back();
        b->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

template <class... Args>
void launch(sp_backend* b, hipFunction_t f, unsigned grid, Args... args) {
    void* params[] = {static_cast<void*>(&args)...};
    if (!b->profiling) {
        hip_check(hipModuleLaunchKernel(f, grid, 1, 1, kWG, 1, 1, 0, b->stream, params, nullptr),
                  "hipModuleLaunchKernel");
        return;
    }
    int id = 0;
    while (id < (int)b->kernel_fns.size() && b->kernel_fns[(size_t)id] != f) ++id;
    sp_backend::Pending pe{id, take_event(b), take_event(b)};
    hip_check(hipEventRecord(pe.start, b->stream), "hipEventRecord");
    hip_check(hipModuleLaunchKernel(f, grid, 1, 1, kWG, 1, 1, 0, b->stream, params, nullptr),
              "hipModuleLaunchKernel");
    hip_check(hipEventRecord(pe.stop, b->stream), "hipEventRecord");
    b->pending.push_back(pe);
    if (b->pending.size() >= 4096) profile_collect(b);
}

unsigned ew_grid(int64_t n) {
    int64_t g = (n + kWG - 1) / kWG;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

// The main kernel of an SpMV without column panels: `k` one workgroup per
// work item, or, for a matrix with a hot table and a kernel that has a
// variant with one, that variant on a persistent grid (as many workgroups as
// the CUs hold at once; each takes the items gridDim apart).
template <class... E>
void spmv_main(sp_backend* b, hipFunction_t k, const DevCsr& a, const void* va, const void* xin, E... epi) {
    hipFunction_t hk = nullptr;
    int occ = 0;
    if (k == b->k_spmv_t_primal_halpern && a.hot64) {
        hk = b->k_spmv_t_primal_halpern_hot;
        occ = b->hot_occ64;
    } else if (k == b->k_spmv_t_primal_mixed && a.hot32) {
        hk = b->k_spmv_t_primal_mixed_hot;
        occ = b->hot_occ32;
    }
    if (hk && occ > 0) {
        const int grid = (int)std::min<int64_t>(a.nblocks, (int64_t)occ * std::max(b->cus, 1));
        launch(b, hk, (unsigned)grid, a.desc, a.rp, a.ci, va, xin, (void*)a.partials, a.nblocks, epi...);
        return;
    }
    launch(b, k, (unsigned)a.nblocks, a.desc, a.rp, a.ci, va, xin, (void*)a.partials, epi...);
}

// One SpMV with a fused epilogue: the SpMV kernel, then, if the matrix has
// rows of several chunks, the finishing kernel with the same epilogue
// arguments.
template <class... E>
void spmv_launch(sp_backend* b, hipFunction_t k, hipFunction_t kfin, const DevCsr& a, const void* va,
                 const void* xin, E... epi) {
    if (a.pmat) {
        // Column panels: the panel matrix's sums (panel by panel), the short
        // rows, then each long row's panel sums added in panel order with the
        // epilogue. FP32 values in, FP32 panel sums out.
        const DevCsr& B = *a.pmat;
        const bool f32 = a.va32 && va == (const void*)a.va32;
        const void* bva = f32 ? (const void*)B.va32 : (const void*)B.va;
        hipFunction_t kp = f32 ? b->k_spmv_panel32 : b->k_spmv_panel;
        hipFunction_t kpf = f32 ? b->k_fin_panel32 : b->k_fin_panel;
        if (B.nblocks) launch(b, kp, (unsigned)B.nblocks, B.desc, B.rp, B.ci, bva, xin, (void*)B.partials, (void*)a.psum);
        if (B.nlong)
            launch(b, kpf, (unsigned)((B.nlong + 3) / 4), (const LongRow*)B.longs, B.nlong, (const void*)B.partials,
                   (void*)a.psum);
        if (a.npdesc) launch(b, k, (unsigned)a.npdesc, a.pdesc, a.rp, a.ci, va, xin, (void*)a.partials, epi...);
        launch(b, kfin, (unsigned)((a.plong + 3) / 4), (const LongRow*)a.plongs, a.plong, (const void*)a.psum, epi...);
        return;
    }
    if (a.nblocks) spmv_main(b, k, a, va, xin, epi...);
    if (a.nlong)
        launch(b, kfin, (unsigned)((a.nlong + 3) / 4), (const LongRow*)a.longs, a.nlong, (const void*)a.partials,
               epi...);
}

void need_len(const sp_vec* v, int64_t len, const char* what) {
    if (!v) throw Error(SP_ERR_INVALID, std::string(what) + ": null vector");
    if (v->len != len)
        throw Error(SP_ERR_SHAPE, std::string(what) + ": length " + std::to_string(v->len) + ", expected " +
                                      std::to_string(len));
}

void plan_csr(sp_backend* b, DevCsr& d, const int64_t* row_ptr, const SpmvTuning& tuning);

// Uploads one CSR matrix with int32 indices and builds its SpMV plan.
void upload_csr(sp_backend* b, DevCsr& d, const int64_t* row_ptr, const int64_t* col_idx, const double* values,
                int64_t nrows, int64_t ncols, int64_t nnz, const SpmvTuning& tuning) {
    if (nrows > INT_MAX - 1 || ncols > INT_MAX - 1 || nnz > INT_MAX)
        throw Error(SP_ERR_TOO_LARGE, "problem too large for 32-bit sparse indices");
    d.nrows = (int)nrows;
    d.ncols = (int)ncols;
    d.nnz = nnz;
    std::vector<int> rp((size_t)nrows + 1), ci((size_t)nnz);
    for (int64_t i = 0; i <= nrows; ++i) rp[(size_t)i] = (int)row_ptr[i];
    for (int64_t k = 0; k < nnz; ++k) ci[(size_t)k] = (int)col_idx[k];
    d.rp = dupload(b->stream, rp.data(), rp.size());
    d.ci = dupload(b->stream, ci.data(), ci.size());
    d.va = dupload(b->stream, values, (size_t)nnz);
    plan_csr(b, d, row_ptr, tuning);
}

// The SpMV plan of a matrix whose arrays are on the device.
void plan_csr(sp_backend* b, DevCsr& d, const int64_t* row_ptr, const SpmvTuning& tuning) {
    const int64_t nrows = d.nrows;
    d.tuning = tuning;
    SpmvPlan plan = plan_spmv(row_ptr, nrows, d.tuning);
    d.nblocks = (int)plan.desc.size();
    d.desc = dupload(b->stream, plan.desc.data(), plan.desc.size());
    d.npartials = plan.npartials;
    d.partials = dmalloc<double>((size_t)std::max(1, d.npartials));
    d.nlong = (int)plan.longs.size();
    d.longs = dupload(b->stream, plan.longs.data(), plan.longs.size());
    d.host_desc = std::move(plan.desc);
    d.host_longs = std::move(plan.longs);
    // The host staging vectors die at return: finish the copies first.
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
}

void free_csr(DevCsr& d);

void free_panels(DevCsr& d) {
    if (d.pmat) {
        free_csr(*d.pmat);
        delete d.pmat;
    }
    dfree(d.psum), dfree(d.plongs), dfree(d.pdesc);
    d.pmat = nullptr, d.psum = nullptr, d.plongs = nullptr, d.pdesc = nullptr;
    d.plong = d.npanels = d.pshift = d.npdesc = 0;
}

void free_csr(DevCsr& d) {
    free_panels(d);
    dfree(d.rp), dfree(d.ci), dfree(d.va), dfree(d.va32), dfree(d.desc), dfree(d.partials), dfree(d.longs);
    d = DevCsr{};
}

}  // namespace

SpmvPlan plan_spmv(const int64_t* rp, int64_t nrows, const SpmvTuning& t) {
    SpmvPlan plan;
    std::vector<Desc> longs, stream;
    int np = 0;
    const int64_t block_nnz = std::max<int64_t>(1, std::min<int64_t>(t.block_nnz, kStreamNnzF64));
    const int64_t long_row = std::max<int64_t>(1, std::min<int64_t>(t.long_row, block_nnz));
    const int64_t max_rows = std::max<int64_t>(1, std::min<int64_t>(t.max_rows, kStreamRows));
    int64_t r = 0;
    while (r < nrows) {
        const int64_t len = rp[r + 1] - rp[r];
        if (len > long_row) {
            // Workgroups of its own, one per chunk; a row of several chunks
            // is finished by the finishing kernel.
            const int64_t nch = (len + kChunk - 1) / kChunk;
            if (nch == 1) {
                longs.push_back(Desc{(int)r, -1, (int)rp[r], (int)rp[r + 1]});
            } else {
                for (int64_t c = 0; c < nch; ++c) {
                    const int64_t k0 = rp[r] + c * kChunk, k1 = std::min(rp[r] + (c + 1) * kChunk, rp[r + 1]);
                    longs.push_back(Desc{(int)r, -1 - (np + (int)c), (int)k0, (int)k1});
                }
                plan.longs.push_back(LongRow{(int)r, np, (int)nch, 0});
                np += (int)nch;
            }
            ++r;
            continue;
        }
        // A stream block: consecutive rows while the nonzeros and rows fit
        // (the merge-path reduction balances any mix of row lengths).
        const int64_t start = r, k0 = rp[r];
        while (r < nrows && rp[r + 1] - rp[r] <= long_row && rp[r + 1] - k0 <= block_nnz && r - start < max_rows)
            ++r;
        stream.push_back(Desc{(int)start, (int)r, (int)k0, (int)rp[r]});
    }
    // Long rows first: they take the longest, so they are dispatched first.
    plan.desc = std::move(longs);
    plan.desc.insert(plan.desc.end(), stream.begin(), stream.end());
    plan.npartials = np;
    return plan;
}

void profile_collect(sp_backend* b) {
    if (b->pending.empty()) return;
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    for (const auto& pe : b->pending) {
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, pe.start, pe.stop), "hipEventElapsedTime");
        if (pe.kernel < (int)b->kernel_ms.size()) {
            b->kernel_ms[(size_t)pe.kernel] += ms;
            b->kernel_launches[(size_t)pe.kernel] += 1;
        }
        b->event_pool.push_back(pe.start);
        b->event_pool.push_back(pe.stop);
    }
    b->pending.clear();
}

void free_problem(sp_problem* p) {
    if (!p) return;
    free_csr(p->a);
    free_csr(p->at);
    for (double* v : {p->c, p->lc, p->uc, p->lv, p->uv, p->pf, p->df}) dfree(v);
    delete p;
}

void free_csr_public(DevCsr& d) { free_csr(d); }

void free_vec32(sp_vec32* v) {
    if (!v) return;
    dfree(v->d);
    delete v;
}

void free_vec(sp_vec* v) {
    if (!v) return;
    dfree(v->d);
    delete v;
}

sp_problem* Hip::load(const ScaledProblem& s, bool autotune) {
    sp_scaled_view v{};
    v.a = {s.a.nrows, s.a.ncols, s.a.nnz(), s.a.row_ptr.data(), s.a.col_idx.data(), s.a.values.data()};
    v.at = {s.at.nrows, s.at.ncols, s.at.nnz(), s.at.row_ptr.data(), s.at.col_idx.data(), s.at.values.data()};
    v.c = s.c.data();
    v.lc = s.lc.data();
    v.uc = s.uc.data();
    v.lv = s.lv.data();
    v.uv = s.uv.data();
    v.row_scale = s.row_scale.data();
    v.col_scale = s.col_scale.data();
    v.bound_scale = s.bound_scale;
    v.obj_scale = s.obj_scale;
    if (!autotune) return load_view(v);
    const SpmvTuning ta = autotuned_plan(s.a.row_ptr.data(), s.a.nrows);
    const SpmvTuning tat = autotuned_plan(s.at.row_ptr.data(), s.at.nrows);
    return load_view(v, &ta, &tat);
}

sp_problem* Hip::load_scaled(const LpView& p_in, const ScalingParams& sp, bool autotune, ScaledProblem& s,
                             ColumnOrder* order) {
    LpView p = p_in;
    RelabeledLp relabeled;  // c, lv, uv in the new column order (when relabeling)
    const int64_t* rank = nullptr;
    const int64_t m = p.m, n = p.n;
    const SpmvTuning ta = autotune ? autotuned_plan(p.row_ptr, m) : SpmvTuning{};
    auto* prob = new sp_problem;
    prob->b = b;
    prob->m = m;
    prob->n = n;
    std::vector<double*> scratch;
    auto tmp = [&](int64_t len) {
        scratch.push_back(dmalloc<double>((size_t)std::max<int64_t>(len, 1)));
        return scratch.back();
    };
    try {
        upload_csr(b, prob->a, p.row_ptr, p.col_idx, p.values, m, n, p.nnz, ta);
        if (order) {
            // Column counts on the device, the order on the host (O(n)), then
            // the column indices relabeled on the device.
            std::vector<int> c32((size_t)n, 0);
            int* dc = dupload(b->stream, c32.data(), c32.size());
            if (p.nnz) launch(b, b->k_col_count, ew_grid(p.nnz), (const int*)prob->a.ci, (long long)p.nnz, dc);
            hip_check(hipMemcpyAsync(c32.data(), dc, 4 * (size_t)n, hipMemcpyDeviceToHost, b->stream), "D2H");
            hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
            *order = popularity_order_of_counts(std::vector<int64_t>(c32.begin(), c32.end()));
            for (int64_t j = 0; j < n; ++j) c32[(size_t)j] = (int)order->rank[(size_t)j];
            hip_check(hipMemcpyAsync(dc, c32.data(), 4 * (size_t)n, hipMemcpyHostToDevice, b->stream), "H2D");
            if (p.nnz) launch(b, b->k_remap, ew_grid(p.nnz), prob->a.ci, (const int*)dc, (long long)p.nnz);
            hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
            dfree(dc);
            relabeled = relabel_vectors(p_in, *order);
            p = relabeled.v;
            rank = order->rank.data();
        }
        std::vector<int64_t> trp;
        transpose(prob->a, prob->at, autotune, trp);
        DevCsr& A = prob->a;
        DevCsr& T = prob->at;
        double *rs = tmp(m), *cs = tmp(n), *r1 = tmp(m), *r2 = tmp(m), *c1 = tmp(n), *c2 = tmp(n), *rf = tmp(m),
               *cf = tmp(n);
        const double* none = nullptr;
        if (m) launch(b, b->k_fill, ew_grid(m), rs, 1.0, (long long)m);
        if (n) launch(b, b->k_fill, ew_grid(n), cs, 1.0, (long long)n);
        auto wave_grid = [](int64_t rows) { return (unsigned)std::max<int64_t>(1, (rows * 64 + kWG - 1) / kWG); };
        auto minmax = [&](const DevCsr& M, double* lo, double* hi) {
            if (M.nrows) launch(b, b->k_sc_minmax, wave_grid(M.nrows), (const int*)M.rp, (const double*)M.va, M.nrows, lo, hi);
        };
        auto powsum = [&](const DevCsr& M, double e, double* out) {
            if (M.nrows)
                launch(b, b->k_sc_powsum, (unsigned)((M.nrows + kWG - 1) / kWG), (const int*)M.rp, (const double*)M.va,
                       M.nrows, e, out);
        };
        auto factor = [&](const double* a, const double* bb, double* f, int64_t len) {
            if (len) launch(b, b->k_sc_factor, ew_grid(len), a, bb, f, (long long)len);
        };
        // v ← v·(rowf·colf) on A and, with the roles swapped, on A'ᵀ.
        auto apply = [&](const double* r, const double* c) {
            if (A.nrows) launch(b, b->k_sc_apply, wave_grid(A.nrows), (const int*)A.rp, (const int*)A.ci, A.va, A.nrows, r, c);
            if (T.nrows) launch(b, b->k_sc_apply, wave_grid(T.nrows), (const int*)T.rp, (const int*)T.ci, T.va, T.nrows, c, r);
        };
        auto mul = [&](double* v, const double* f, int64_t len) {
            if (len) launch(b, b->k_sc_mul, ew_grid(len), v, f, (long long)len);
        };
        // scaling.rs:59-70: geometric-mean passes, rows then columns.
        for (int64_t pass = 0; pass < sp.geometric_mean_passes; ++pass) {
            minmax(A, r1, r2);
            factor(r1, r2, rf, m);
            apply(rf, none);
            mul(rs, rf, m);
            minmax(T, c1, c2);
            factor(c1, c2, cf, n);
            apply(none, cf);
            mul(cs, cf, n);
        }
        // scaling.rs:71-76: Ruiz by the square root of the line maxima.
        for (int64_t it = 0; it < sp.ruiz_iterations; ++it) {
            minmax(A, nullptr, r2);
            minmax(T, nullptr, c2);
            factor(r2, none, rf, m);
            factor(c2, none, cf, n);
            apply(rf, cf);
            mul(rs, rf, m);
            mul(cs, cf, n);
        }
        // scaling.rs:77-84: Pock-Chambolle.
        if (sp.has_pock_chambolle) {
            powsum(A, sp.pock_chambolle_alpha, r1);
            factor(r1, none, rf, m);
            powsum(T, 2.0 - sp.pock_chambolle_alpha, c1);
            factor(c1, none, cf, n);
            apply(rf, cf);
            mul(rs, rf, m);
            mul(cs, cf, n);
        }
        s.row_scale.resize((size_t)m);
        s.col_scale.resize((size_t)n);
        if (m) hip_check(hipMemcpyAsync(s.row_scale.data(), rs, 8 * (size_t)m, hipMemcpyDeviceToHost, b->stream), "D2H");
        if (n) hip_check(hipMemcpyAsync(s.col_scale.data(), cs, 8 * (size_t)n, hipMemcpyDeviceToHost, b->stream), "D2H");
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        for (double* q : scratch) dfree(q);
        scratch.clear();
        // The vector part on the host, as ScaledProblem::new does it.
        s.a = Csr{};
        s.at = Csr{};
        s.a.nrows = m, s.a.ncols = n, s.at.nrows = n, s.at.ncols = m;
        finish_scaling(p, sp, s, rank);
        std::vector<double> pf((size_t)m), df((size_t)n);
        for (int64_t i = 0; i < m; ++i) pf[(size_t)i] = 1.0 / (s.row_scale[(size_t)i] * s.bound_scale);
        for (int64_t j = 0; j < n; ++j) df[(size_t)j] = 1.0 / (s.col_scale[(size_t)j] * s.obj_scale);
        prob->c = dupload(b->stream, s.c.data(), (size_t)n);
        prob->lv = dupload(b->stream, s.lv.data(), (size_t)n);
        prob->uv = dupload(b->stream, s.uv.data(), (size_t)n);
        prob->lc = dupload(b->stream, s.lc.data(), (size_t)m);
        prob->uc = dupload(b->stream, s.uc.data(), (size_t)m);
        prob->pf = dupload(b->stream, pf.data(), (size_t)m);
        prob->df = dupload(b->stream, df.data(), (size_t)n);
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    } catch (...) {
        (void)hipStreamSynchronize(b->stream);
        for (double* q : scratch) dfree(q);
        free_problem(prob);
        throw;
    }
    return prob;
}

namespace {
void scan_ints(sp_backend* b, int* a, int64_t len, std::vector<int*>& scratch);
}  // namespace

void Hip::set_panels(sp_problem* prob, int shift) {
    if (!prob) throw Error(SP_ERR_INVALID, "set_panels: null problem");
    DevCsr& A = prob->a;
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    free_panels(A);
    if (shift == 0) return;
    if (shift < 8 || shift > 30) throw Error(SP_ERR_INVALID, "set_panels: panel width 2^shift needs shift in [8, 30]");
    const int m = A.nrows, n = A.ncols;
    std::vector<int> rp32((size_t)m + 1);
    hip_check(hipMemcpyAsync(rp32.data(), A.rp, sizeof(int) * rp32.size(), hipMemcpyDeviceToHost, b->stream),
              "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    // The panel rows are the rows the short rows' plan below leaves out.
    SpmvTuning ts = A.tuning;
    ts.long_row = std::min(ts.block_nnz, kPanelMinRow);
    std::vector<int> rows;
    for (int r = 0; r < m; ++r)
        if (rp32[(size_t)r + 1] - rp32[(size_t)r] > ts.long_row) rows.push_back(r);
    if (rows.empty()) return;
    const int L = (int)rows.size();
    const int64_t P64 = ((int64_t)n + (int64_t(1) << shift) - 1) >> shift;
    const int64_t nq = (int64_t)L * P64;
    if (nq + 1 > INT_MAX / 2) throw Error(SP_ERR_TOO_LARGE, "set_panels: too many (panel, long row) pairs");
    const int P = (int)P64;
    std::vector<int*> scratch;
    auto* B = new DevCsr;
    // The panel parts of a row are contiguous only where its columns are
    // sorted, which relabeling by popularity does not keep. A'ᵀ transposed
    // back has A's rows with their columns in increasing order (its rows are
    // scanned in column order); the panel matrix takes its nonzeros from it.
    DevCsr S;
    try {
        std::vector<int64_t> srp;
        transpose(prob->at, S, false, srp);
        int* drows = dupload(b->stream, rows.data(), rows.size());
        scratch.push_back(drows);
        int* start = dmalloc<int>((size_t)nq);
        scratch.push_back(start);
        B->rp = dmalloc<int>((size_t)nq + 1);  // the counts, then their exclusive scan
        launch(b, b->k_panel_count, (unsigned)((nq + kWG - 1) / kWG), (const int*)S.rp, (const int*)S.ci,
               (const int*)drows, L, P, shift, start, B->rp);
        hip_check(hipMemsetAsync(B->rp + nq, 0, sizeof(int), b->stream), "hipMemsetAsync");
        scan_ints(b, B->rp, nq + 1, scratch);
        std::vector<int> brp32((size_t)nq + 1);
        hip_check(hipMemcpyAsync(brp32.data(), B->rp, sizeof(int) * brp32.size(), hipMemcpyDeviceToHost, b->stream),
                  "hipMemcpyAsync D2H");
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        B->nrows = (int)nq;
        B->ncols = n;
        B->nnz = brp32[(size_t)nq];
        B->ci = dmalloc<int>((size_t)std::max<int64_t>(B->nnz, 1));
        B->va = dmalloc<double>((size_t)std::max<int64_t>(B->nnz, 1));
        launch(b, b->k_panel_gather, (unsigned)((nq * 64 + kWG - 1) / kWG), (const int*)start, (const int*)B->rp,
               (long long)nq, (const int*)S.ci, (const double*)S.va, B->ci, B->va);
        if (A.va32) {
            B->va32 = dmalloc<float>((size_t)std::max<int64_t>(B->nnz, 1));
            if (B->nnz) launch(b, b->k_to_f32, ew_grid(B->nnz), (const double*)B->va, B->va32, (long long)B->nnz);
        }
        const std::vector<int64_t> brp(brp32.begin(), brp32.end());
        plan_csr(b, *B, brp.data(), SpmvTuning{});  // panel-major stream blocks (synchronizes)
        // Each long row's P panel sums, added in panel order by the finishing kernels.
        std::vector<LongRow> lr((size_t)L);
        for (int l = 0; l < L; ++l) lr[(size_t)l] = LongRow{rows[(size_t)l], l, P, L};
        A.plongs = dupload(b->stream, lr.data(), lr.size());
        A.psum = dmalloc<double>((size_t)nq);
        // The short rows: the stream blocks of a plan whose long rows are the panel rows.
        const std::vector<int64_t> rp64(rp32.begin(), rp32.end());
        SpmvPlan sp = plan_spmv(rp64.data(), m, ts);
        std::vector<Desc> shorts;
        for (const Desc& d : sp.desc)
            if (d.r1 >= 0) shorts.push_back(d);
        A.npdesc = (int)shorts.size();
        A.pdesc = shorts.empty() ? dmalloc<Desc>(1) : dupload(b->stream, shorts.data(), shorts.size());
        A.pmat = B;
        A.plong = L;
        A.npanels = P;
        A.pshift = shift;
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    } catch (...) {
        (void)hipStreamSynchronize(b->stream);
        for (int* q : scratch) dfree(q);
        free_csr(S);
        if (A.pmat != B) {
            free_csr(*B);
            delete B;
        }
        free_panels(A);
        throw;
    }
    for (int* q : scratch) dfree(q);
    free_csr(S);
}

void Hip::permute_rows(sp_problem* prob, const std::vector<int64_t>& order) {
    if (!prob) throw Error(SP_ERR_INVALID, "permute_rows: null problem");
    DevCsr& A = prob->a;
    const int m = A.nrows;
    if ((int64_t)order.size() != m) throw Error(SP_ERR_SHAPE, "permute_rows: order has the wrong length");
    if (prob->has_fp32 || A.pmat) throw Error(SP_ERR_INVALID, "permute_rows: after FP32 copies or panels");
    if (m == 0) return;  // nothing to relabel
    std::vector<int> ord32((size_t)m), rank((size_t)m, -1);
    for (int i = 0; i < m; ++i) {
        const int64_t o = order[(size_t)i];
        if (o < 0 || o >= m || rank[(size_t)o] >= 0) throw Error(SP_ERR_INVALID, "permute_rows: not a permutation");
        ord32[(size_t)i] = (int)o;
        rank[(size_t)o] = i;
    }
    std::vector<int> rp_old((size_t)m + 1);
    hip_check(hipMemcpyAsync(rp_old.data(), A.rp, sizeof(int) * rp_old.size(), hipMemcpyDeviceToHost, b->stream),
              "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    std::vector<int64_t> rp_new((size_t)m + 1, 0);
    for (int i = 0; i < m; ++i)
        rp_new[(size_t)i + 1] = rp_new[(size_t)i] + rp_old[(size_t)ord32[(size_t)i] + 1] - rp_old[(size_t)ord32[(size_t)i]];
    std::vector<int> rp_new32(rp_new.begin(), rp_new.end());
    std::vector<void*> scratch;
    int *drp = nullptr, *dci = nullptr;
    double* dva = nullptr;
    try {
        int* dord = dupload(b->stream, ord32.data(), std::max<size_t>(ord32.size(), 1));
        scratch.push_back(dord);
        int* drank = dupload(b->stream, rank.data(), std::max<size_t>(rank.size(), 1));
        scratch.push_back(drank);
        drp = dupload(b->stream, rp_new32.data(), rp_new32.size());
        dci = dmalloc<int>((size_t)std::max<int64_t>(A.nnz, 1));
        dva = dmalloc<double>((size_t)std::max<int64_t>(A.nnz, 1));
        if (m)
            launch(b, b->k_rows_gather, (unsigned)(((int64_t)m * 64 + kWG - 1) / kWG), (const int*)A.rp,
                   (const int*)A.ci, (const double*)A.va, (const int*)drp, (const int*)dord, m, dci, dva);
        // The row-indexed vectors, and A'ᵀ's column indices (its rows keep their order).
        for (double** v : {&prob->lc, &prob->uc, &prob->pf}) {
            double* nv = dmalloc<double>((size_t)std::max(m, 1));
            if (m) launch(b, b->k_gather, ew_grid(m), (const double*)*v, (const int*)dord, nv, (long long)m);
            hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
            dfree(*v);
            *v = nv;
        }
        if (prob->at.nnz) launch(b, b->k_remap, ew_grid(prob->at.nnz), prob->at.ci, (const int*)drank, (long long)prob->at.nnz);
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    } catch (...) {
        (void)hipStreamSynchronize(b->stream);
        for (void* q : scratch) dfree(q);
        dfree(drp), dfree(dci), dfree(dva);
        throw;
    }
    for (void* q : scratch) dfree(q);
    // A' takes the new arrays and a plan for its new row order.
    const SpmvTuning t = A.tuning;
    dfree(A.rp), dfree(A.ci), dfree(A.va), dfree(A.desc), dfree(A.partials), dfree(A.longs);
    A.rp = drp, A.ci = dci, A.va = dva;
    A.desc = nullptr, A.partials = nullptr, A.longs = nullptr;
    plan_csr(b, A, rp_new.data(), t);
}

void Hip::plan_hot(sp_problem* prob, int mode) {
    if (!prob) throw Error(SP_ERR_INVALID, "plan_hot: null problem");
    if (mode < -1 || mode > 1) throw Error(SP_ERR_INVALID, "hot_table must be -1, 0 or 1");
    DevCsr& T = prob->at;
    const DevCsr& A = prob->a;
    T.hot64 = T.hot32 = false;
    // Worth a persistent grid and its LDS copies only on a large matrix (each
    // workgroup's copy reads 16 KB: about 12-16 MB per launch, 1% of a
    // 100M-nonzero pass) and only when most gathers hit the table.
    if (mode == 0 || T.pmat || T.nblocks == 0 || (mode < 0 && T.nnz < kHotMinNnz)) return;
    auto frac = [&](int h) {
        if (A.nrows < h || T.ncols < h) return 0.0;
        int cnt = 0;
        hip_check(hipMemcpyAsync(&cnt, A.rp + h, sizeof(int), hipMemcpyDeviceToHost, b->stream), "hipMemcpyAsync D2H");
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        return (double)cnt / (double)T.nnz;
    };
    T.hot64 = frac(kHotF64) >= kHotMinFrac;
    T.hot32 = frac(kHotF32) >= kHotMinFrac;
}

void Hip::upload(DevCsr& d, const int64_t* row_ptr, const int64_t* col_idx, const double* values, int64_t nrows,
                 int64_t ncols, int64_t nnz, const SpmvTuning& tuning) {
    upload_csr(b, d, row_ptr, col_idx, values, nrows, ncols, nnz, tuning);
}

namespace {

// In-place exclusive scan of `len` ints on the device (multi-level).
void scan_ints(sp_backend* b, int* a, int64_t len, std::vector<int*>& scratch) {
    const int64_t per = 4 * kWG;
    const int64_t nb = (len + per - 1) / per;
    if (nb == 0) return;
    scratch.push_back(dmalloc<int>((size_t)nb));
    int* sums = scratch.back();
    launch(b, b->k_scan_tiles, (unsigned)nb, a, (long long)len, sums);
    if (nb > 1) {
        scan_ints(b, sums, nb, scratch);
        launch(b, b->k_scan_add, (unsigned)nb, a, (long long)len, (const int*)sums);
    }
}

constexpr int kTrTile = 256 * 16;  // TR_TILE in kernels.hip

}  // namespace

void Hip::transpose(const DevCsr& A, DevCsr& T, bool autotune, std::vector<int64_t>& host_rp) {
    const int nnz = (int)A.nnz, ncols = A.ncols, nrows = A.nrows;
    T = DevCsr{};
    T.nrows = ncols;
    T.ncols = nrows;
    T.nnz = A.nnz;
    std::vector<int*> scratch;
    auto ialloc = [&](int64_t len) {
        scratch.push_back(dmalloc<int>((size_t)std::max<int64_t>(len, 1)));
        return scratch.back();
    };
    try {
        T.rp = dmalloc<int>((size_t)ncols + 1);
        T.ci = dmalloc<int>((size_t)std::max(nnz, 1));
        T.va = dmalloc<double>((size_t)std::max(nnz, 1));
        int* keys = ialloc(nnz);
        if (nnz > 0) {
            int *rows = ialloc(nnz), *vals = ialloc(nnz), *keys2 = ialloc(nnz), *vals2 = ialloc(nnz);
            const int ntiles = (nnz + kTrTile - 1) / kTrTile;
            int* hist = ialloc((int64_t)256 * ntiles);
            launch(b, b->k_tr_rows, (unsigned)std::max<int64_t>(1, ((int64_t)nrows * 64 + kWG - 1) / kWG),
                   (const int*)A.rp, nrows, rows);
            hip_check(hipMemcpyAsync(keys, A.ci, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToDevice, b->stream),
                      "hipMemcpyAsync D2D");
            launch(b, b->k_tr_iota, ew_grid(nnz), vals, (long long)nnz);
            int bits = 1;
            while (bits < 31 && (int64_t(1) << bits) < ncols) ++bits;
            for (int shift = 0; shift < bits; shift += 8) {
                launch(b, b->k_tr_hist, (unsigned)ntiles, (const int*)keys, nnz, shift, hist, ntiles);
                scan_ints(b, hist, (int64_t)256 * ntiles, scratch);
                launch(b, b->k_tr_scatter, (unsigned)ntiles, (const int*)keys, (const int*)vals, nnz, shift,
                       (const int*)hist, ntiles, keys2, vals2);
                std::swap(keys, keys2);
                std::swap(vals, vals2);
            }
            launch(b, b->k_tr_gather, ew_grid(nnz), (const int*)vals, (const int*)rows, (const double*)A.va,
                   (long long)nnz, T.ci, T.va);
        }
        launch(b, b->k_tr_ptr, (unsigned)((int64_t(nnz) + 1 + kWG - 1) / kWG), (const int*)keys, nnz, ncols, T.rp);
        std::vector<int> rp((size_t)ncols + 1);
        hip_check(hipMemcpyAsync(rp.data(), T.rp, sizeof(int) * rp.size(), hipMemcpyDeviceToHost, b->stream),
                  "hipMemcpyAsync D2H");
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        for (int* q : scratch) dfree(q);
        scratch.clear();
        host_rp.assign(rp.begin(), rp.end());
        plan_csr(b, T, host_rp.data(), autotune ? autotuned_plan(host_rp.data(), ncols) : SpmvTuning{});
    } catch (...) {
        (void)hipStreamSynchronize(b->stream);
        for (int* q : scratch) dfree(q);
        free_csr(T);
        throw;
    }
}

sp_kkt_report Hip::check(const LpView& p, const double* x, const double* y) {
    const int64_t m = p.m, n = p.n;
    if (!x || !y) throw Error(SP_ERR_INVALID, "check: null vector");
    DevCsr A, T;
    std::vector<double*> vecs;
    auto up = [&](const double* h, int64_t len) {
        vecs.push_back(dupload(b->stream, h, (size_t)std::max<int64_t>(len, 0)));
        return vecs.back();
    };
    double r[7] = {0, 0, 0, 0, 0, 0, 0};
    try {
        upload_csr(b, A, p.row_ptr, p.col_idx, p.values, m, n, p.nnz, SpmvTuning{});
        std::vector<int64_t> trp;
        transpose(A, T, false, trp);
        double *dx = up(x, n), *dy = up(y, m), *c = up(p.c, n), *lc = up(p.lc, m), *uc = up(p.uc, m),
               *lv = up(p.lv, n), *uv = up(p.uv, n);
        vecs.push_back(dmalloc<double>((size_t)std::max<int64_t>(m, 1)));
        double* ax = vecs.back();
        vecs.push_back(dmalloc<double>((size_t)std::max<int64_t>(n, 1)));
        double* aty = vecs.back();
        spmv_launch(b, b->k_spmv, b->k_fin, A, A.va, dx, ax);
        spmv_launch(b, b->k_spmv, b->k_fin, T, T.va, dy, aty);
        const int grid = reduce_grid(m + n);
        launch(b, b->k_red_check, (unsigned)grid, (const double*)ax, (const double*)dy, (const double*)lc,
               (const double*)uc, (long long)m, (const double*)dx, (const double*)aty, (const double*)c,
               (const double*)lv, (const double*)uv, (long long)n, b->red_partials, b->red_result);
        if (grid > 1) launch(b, b->k_red_finish7, 1u, (const double*)b->red_partials, grid, b->red_result);
        hip_check(hipMemcpyAsync(r, b->red_result, sizeof r, hipMemcpyDeviceToHost, b->stream), "D2H");
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    } catch (...) {
        (void)hipStreamSynchronize(b->stream);
        free_csr(A);
        free_csr(T);
        for (double* v : vecs) dfree(v);
        throw;
    }
    free_csr(A);
    free_csr(T);
    for (double* v : vecs) dfree(v);
    // check/src/lib.rs:76-117 on the sums.
    sp_kkt_report k;
    k.primal_objective = r[0] + p.c0;
    k.dual_objective = r[3] + p.c0;
    k.constraint_violation = std::sqrt(r[1]);
    k.bound_violation = std::sqrt(r[2]);
    k.dual_violation = std::sqrt(r[4]);
    const double b_norm = std::sqrt(r[5]), c_norm = std::sqrt(r[6]);
    const double primal = std::sqrt(k.constraint_violation * k.constraint_violation +
                                    k.bound_violation * k.bound_violation);
    k.relative_primal = primal / (1.0 + b_norm);
    k.relative_dual = k.dual_violation / (1.0 + c_norm);
    k.relative_gap = std::fabs(k.primal_objective - k.dual_objective) /
                     (1.0 + std::fabs(k.primal_objective) + std::fabs(k.dual_objective));
    return k;
}

sp_problem* Hip::load_view(const sp_scaled_view& v, const SpmvTuning* ta, const SpmvTuning* tat) {
    const int64_t m = v.a.nrows, n = v.a.ncols;
    if (v.at.nrows != n || v.at.ncols != m || v.at.nnz != v.a.nnz)
        throw Error(SP_ERR_SHAPE, "load: A'ᵀ does not match A'");
    auto* p = new sp_problem;
    p->b = b;
    p->m = m;
    p->n = n;
    try {
        upload_csr(b, p->a, v.a.row_ptr, v.a.col_idx, v.a.values, m, n, v.a.nnz, ta ? *ta : SpmvTuning{});
        upload_csr(b, p->at, v.at.row_ptr, v.at.col_idx, v.at.values, n, m, v.at.nnz, tat ? *tat : SpmvTuning{});
        // Residual factors are formed on the host with the reference's
        // arithmetic (scaling.rs:124-131) and uploaded, not recomputed.
        std::vector<double> pf((size_t)m), df((size_t)n);
        for (int64_t i = 0; i < m; ++i) pf[(size_t)i] = 1.0 / (v.row_scale[i] * v.bound_scale);
        for (int64_t j = 0; j < n; ++j) df[(size_t)j] = 1.0 / (v.col_scale[j] * v.obj_scale);
        p->c = dupload(b->stream, v.c, (size_t)n);
        p->lv = dupload(b->stream, v.lv, (size_t)n);
        p->uv = dupload(b->stream, v.uv, (size_t)n);
        p->lc = dupload(b->stream, v.lc, (size_t)m);
        p->uc = dupload(b->stream, v.uc, (size_t)m);
        p->pf = dupload(b->stream, pf.data(), (size_t)m);
        p->df = dupload(b->stream, df.data(), (size_t)n);
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    } catch (...) {
        free_problem(p);
        throw;
    }
    return p;
}

sp_vec* Hip::upload(const double* data, int64_t len) {
    if (len < 0) throw Error(SP_ERR_INVALID, "upload: negative length");
    auto* v = new sp_vec;
    v->b = b;
    v->len = len;
    try {
        v->d = dmalloc<double>((size_t)len);
        if (len) {
            hip_check(hipMemcpyAsync(v->d, data, sizeof(double) * (size_t)len, hipMemcpyHostToDevice, b->stream),
                      "hipMemcpyAsync H2D");
            hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        }
    } catch (...) {
        free_vec(v);
        throw;
    }
    return v;
}

sp_vec* Hip::zeros(int64_t len) {
    if (len < 0) throw Error(SP_ERR_INVALID, "zeros: negative length");
    auto* v = new sp_vec;
    v->b = b;
    v->len = len;
    try {
        v->d = dmalloc<double>((size_t)len);
        if (len) launch(b, b->k_fill, ew_grid(len), v->d, 0.0, (long long)len);
    } catch (...) {
        free_vec(v);
        throw;
    }
    return v;
}

void Hip::download(const sp_vec* v, double* out, int64_t len) {
    need_len(v, len, "download");
    if (len) {
        hip_check(hipMemcpyAsync(out, v->d, sizeof(double) * (size_t)len, hipMemcpyDeviceToHost, b->stream),
                  "hipMemcpyAsync D2H");
    }
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
}

void Hip::copy(const sp_vec* src, sp_vec* dst) {
    need_len(dst, src ? src->len : -1, "copy");
    if (src->len)
        hip_check(hipMemcpyAsync(dst->d, src->d, sizeof(double) * (size_t)src->len, hipMemcpyDeviceToDevice,
                                 b->stream),
                  "hipMemcpyAsync D2D");
}

// ---- mixed precision -------------------------------------------------------------------

void Hip::ensure_fp32(sp_problem* p) {
    if (p->has_fp32) return;
    for (DevCsr* a : {&p->a, &p->at, p->a.pmat}) {
        if (!a) continue;
        a->va32 = dmalloc<float>((size_t)a->nnz);
        if (a->nnz) launch(b, b->k_to_f32, ew_grid(a->nnz), (const double*)a->va, a->va32, (long long)a->nnz);
    }
    p->has_fp32 = true;
}

sp_vec32* Hip::zeros32(int64_t len) {
    auto* v = new sp_vec32;
    v->len = len;
    try {
        v->d = dmalloc<float>((size_t)len);
        hip_check(hipMemsetAsync(v->d, 0, sizeof(float) * (size_t)std::max<int64_t>(len, 1), b->stream),
                  "hipMemsetAsync");
    } catch (...) {
        free_vec32(v);
        throw;
    }
    return v;
}

void Hip::delta_from(const sp_vec* z, const sp_vec* anchor, sp_vec32* dz) {
    if (!z || !anchor || !dz || z->len != dz->len || anchor->len != dz->len)
        throw Error(SP_ERR_SHAPE, "delta_from: length mismatch");
    if (z->len)
        launch(b, b->k_delta_from, ew_grid(z->len), (const double*)z->d, (const double*)anchor->d, dz->d,
               (long long)z->len);
}

void Hip::delta_to(const sp_vec32* dz, const sp_vec* anchor, sp_vec* z) {
    if (!z || !anchor || !dz || z->len != dz->len || anchor->len != dz->len)
        throw Error(SP_ERR_SHAPE, "delta_to: length mismatch");
    if (z->len)
        launch(b, b->k_delta_to, ew_grid(z->len), (const float*)dz->d, (const double*)anchor->d, z->d,
               (long long)z->len);
}

void Hip::plain_iteration_mixed(const sp_problem* p, sp_vec32* dx, sp_vec32* dy, const sp_vec* x_base,
                                const sp_vec* y_base, const sp_vec* atya, const sp_vec* axa, const sp_vec32* ex,
                                const sp_vec32* ey, sp_vec32* dx_bar, double tau, double sigma, double w, double rho) {
    if (!p->has_fp32) throw Error(SP_ERR_INVALID, "plain_iteration_mixed: FP32 values not prepared");
    const DevCsr& at = p->at;
    const DevCsr& a = p->a;
    spmv_launch(b, b->k_spmv_t_primal_mixed, b->k_fin_t_primal_mixed, at, at.va32, dy->d, dx->d,
                (const double*)x_base->d, (const double*)p->c, (const double*)p->lv, (const double*)p->uv,
                (const double*)atya->d, (const float*)ex->d, tau, w, rho, dx_bar->d);
    spmv_launch(b, b->k_spmv_dual_mixed, b->k_fin_dual_mixed, a, a.va32, dx_bar->d, dy->d,
                (const double*)y_base->d, (const double*)p->lc, (const double*)p->uc, (const double*)axa->d,
                (const float*)ey->d, sigma, w, rho);
}

// ---- bandwidth probe ---------------------------------------------------------------------

double Hip::probe(int kind, int64_t n, int64_t table_len, int reps, uint64_t seed) {
    if (n <= 0 || reps <= 0 || ((kind == 2 || (kind >= 5 && kind <= 8)) && table_len <= 0))
        throw Error(SP_ERR_INVALID, "probe: bad size");
    if (kind < 0 || kind > 8) throw Error(SP_ERR_INVALID, "probe: unknown kind");
    if (kind >= 5) return probe_overlap(kind - 5, n, table_len, reps, seed);
    if (kind == 3 || kind == 4) {
        // The float copy and triad run on the doubles' buffers, reinterpreted.
        const int64_t half = (n + 1) / 2;
        double *a = nullptr, *bb = nullptr, *c = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        float ms = 0.0f;
        hipError_t err = hipSuccess;
        try {
            a = dmalloc<double>((size_t)half), bb = dmalloc<double>((size_t)half), c = dmalloc<double>((size_t)half);
            launch(b, b->k_fill, ew_grid(half), a, 1.0, (long long)half);
            launch(b, b->k_fill, ew_grid(half), bb, 2.0, (long long)half);
            const unsigned grid = (unsigned)std::min<int64_t>((n + kWG - 1) / kWG, 65536);
            auto once = [&] {
                if (kind == 3) launch(b, b->k_probe_copy32, grid, (const float*)a, (float*)c, (long long)n);
                else launch(b, b->k_probe_triad32, grid, (const float*)a, (const float*)bb, (float*)c, 3.0f, (long long)n);
            };
            once();
            hip_check(hipEventCreate(&e0), "hipEventCreate");
            hip_check(hipEventCreate(&e1), "hipEventCreate");
            err = hipEventRecord(e0, b->stream);
            for (int r = 0; r < reps && err == hipSuccess; ++r) once();
            if (err == hipSuccess) err = hipEventRecord(e1, b->stream);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        } catch (...) {
            (void)hipStreamSynchronize(b->stream);
            for (void* q : {(void*)a, (void*)bb, (void*)c}) dfree(q);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipStreamSynchronize(b->stream);
        for (void* q : {(void*)a, (void*)bb, (void*)c}) dfree(q);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        hip_check(err, "probe events");
        return (double)ms / reps;
    }
    double *a = nullptr, *bb = nullptr, *c = nullptr, *table = nullptr;
    int* idx = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.0f;
    try {
        a = dmalloc<double>((size_t)n);
        c = dmalloc<double>((size_t)n);
        launch(b, b->k_fill, ew_grid(n), a, 1.0, (long long)n);
        if (kind == 1) {
            bb = dmalloc<double>((size_t)n);
            launch(b, b->k_fill, ew_grid(n), bb, 2.0, (long long)n);
        }
        if (kind == 2) {
            table = dmalloc<double>((size_t)table_len);
            launch(b, b->k_fill, ew_grid(table_len), table, 3.0, (long long)table_len);
            std::vector<int> h((size_t)n);
            uint64_t st = seed * 0x9E3779B97F4A7C15ull | 1;
            for (auto& v : h) {
                st ^= st << 13, st ^= st >> 7, st ^= st << 17;
                v = (int)(st % (uint64_t)table_len);
            }
            idx = dupload(b->stream, h.data(), h.size());
            hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        }
        const unsigned grid = (unsigned)std::min<int64_t>((n + kWG - 1) / kWG, 65536);
        auto once = [&] {
            if (kind == 0) launch(b, b->k_probe_copy, grid, (const double*)a, c, (long long)n);
            else if (kind == 1) launch(b, b->k_probe_triad, grid, (const double*)a, (const double*)bb, c, 3.0, (long long)n);
            else launch(b, b->k_probe_gather, grid, (const int*)idx, (const double*)table, c, (long long)n);
        };
        once();
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(hipEventRecord(e0, b->stream), "hipEventRecord");
        for (int r = 0; r < reps; ++r) once();
        hip_check(hipEventRecord(e1, b->stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    } catch (...) {
        (void)hipStreamSynchronize(b->stream);
        for (void* q : {(void*)a, (void*)bb, (void*)c, (void*)table, (void*)idx}) dfree(q);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipStreamSynchronize(b->stream);
    for (void* q : {(void*)a, (void*)bb, (void*)c, (void*)table, (void*)idx}) dfree(q);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / reps;
}

// sp_probe_overlap in mode `mode` over nb = ceil(n / 2048) blocks of each kind.
double Hip::probe_overlap(int mode, int64_t n, int64_t table_len, int reps, uint64_t seed) {
    constexpr int64_t kBlock = 8 * kWG;
    const int64_t nb = (n + kBlock - 1) / kBlock, len = nb * kBlock;
    if (nb > (int64_t(1) << 30)) throw Error(SP_ERR_TOO_LARGE, "probe: too many blocks");
    const int64_t grid = mode == 2 ? 2 * nb : nb;
    double *table = nullptr, *v = nullptr, *out = nullptr;
    int *idx = nullptr, *ci = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.0f;
    try {
        table = dmalloc<double>((size_t)table_len);
        v = dmalloc<double>((size_t)len);
        out = dmalloc<double>((size_t)(grid * kWG));
        launch(b, b->k_fill, ew_grid(table_len), table, 3.0, (long long)table_len);
        launch(b, b->k_fill, ew_grid(len), v, 1.0, (long long)len);
        // Uniform random indices (xorshift); the stream's ints are a second
        // draw, so the two kinds of block share no lines.
        std::vector<int> h((size_t)len);
        uint64_t st = seed * 0x9E3779B97F4A7C15ull | 1;
        auto draw = [&] {
            for (auto& x : h) {
                st ^= st << 13, st ^= st >> 7, st ^= st << 17;
                x = (int)(st % (uint64_t)table_len);
            }
        };
        draw();
        idx = dupload(b->stream, h.data(), h.size());
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        draw();
        ci = dupload(b->stream, h.data(), h.size());
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        auto once = [&] {
            launch(b, b->k_probe_overlap, (unsigned)grid, (const int*)idx, (const double*)table, (const double*)v,
                   (const int*)ci, mode, out);
        };
        once();
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(hipEventRecord(e0, b->stream), "hipEventRecord");
        for (int r = 0; r < reps; ++r) once();
        hip_check(hipEventRecord(e1, b->stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    } catch (...) {
        (void)hipStreamSynchronize(b->stream);
        for (void* q : {(void*)table, (void*)v, (void*)out, (void*)idx, (void*)ci}) dfree(q);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipStreamSynchronize(b->stream);
    for (void* q : {(void*)table, (void*)v, (void*)out, (void*)idx, (void*)ci}) dfree(q);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / reps;
}

void Hip::sync() { hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize"); }

double Hip::spmv_time(const sp_problem* p, bool transpose, int reps) {
    if (!p || reps < 1) throw Error(SP_ERR_INVALID, "spmv_time: bad argument");
    const DevCsr& M = transpose ? p->at : p->a;
    std::unique_ptr<sp_vec, void (*)(sp_vec*)> x(zeros(M.ncols), free_vec), out(zeros(M.nrows), free_vec);
    if (M.ncols) launch(b, b->k_fill, ew_grid(M.ncols), x->d, 1.0, (long long)M.ncols);
    spmv_launch(b, b->k_spmv, b->k_fin, M, M.va, x->d, out->d);  // warm-up
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.0f;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    hipError_t err = hipEventRecord(e0, b->stream);
    for (int r = 0; r < reps && err == hipSuccess; ++r) spmv_launch(b, b->k_spmv, b->k_fin, M, M.va, x->d, out->d);
    if (err == hipSuccess) err = hipEventRecord(e1, b->stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    hip_check(err, "spmv_time events");
    return (double)ms / reps;
}

namespace {

// Milliseconds of `reps` calls of `once` between two events on the stream.
template <class F>
double time_reps(sp_backend* b, int reps, F&& once) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.0f;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hipError_t err = hipEventCreate(&e1);
    if (err == hipSuccess) err = hipEventRecord(e0, b->stream);
    try {
        for (int r = 0; r < reps && err == hipSuccess; ++r) once();
    } catch (...) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    if (err == hipSuccess) err = hipEventRecord(e1, b->stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    hip_check(err, "timing events");
    return (double)ms;
}

}  // namespace

double Hip::gather_replay_time(const sp_problem* p, bool transpose, bool fp32, int reps) {
    if (!p || reps < 1) throw Error(SP_ERR_INVALID, "gather_replay_time: bad argument");
    const DevCsr& M = transpose ? p->at : p->a;
    std::unique_ptr<sp_vec, void (*)(sp_vec*)> table(zeros(std::max(M.ncols, 1)), free_vec);
    double* sink = dmalloc<double>(1);
    hipFunction_t k = fp32 ? b->k_replay32 : b->k_replay64;
    // The work items of the matrix's own plan (with column panels: the panel
    // matrix's, then the short rows'), each loading its nonzeros' table entries.
    auto once = [&] {
        if (M.pmat) {
            const DevCsr& B = *M.pmat;
            if (B.nblocks) launch(b, k, (unsigned)B.nblocks, (const Desc*)B.desc, (const int*)B.ci,
                                  (const void*)table->d, sink);
            if (M.npdesc) launch(b, k, (unsigned)M.npdesc, (const Desc*)M.pdesc, (const int*)M.ci,
                                 (const void*)table->d, sink);
        } else if (M.nblocks) {
            launch(b, k, (unsigned)M.nblocks, (const Desc*)M.desc, (const int*)M.ci, (const void*)table->d, sink);
        }
    };
    double ms = 0.0;
    try {
        once();  // warm-up
        ms = time_reps(b, reps, once) / reps;
    } catch (...) {
        dfree(sink);
        throw;
    }
    (void)hipStreamSynchronize(b->stream);
    dfree(sink);
    return ms;
}

double Hip::plain_kernel_time(const sp_problem* p, int side, bool mixed, const PlainArgs& a, int reps) {
    if (!p || reps < 1 || side < 0 || side > 1) throw Error(SP_ERR_INVALID, "plain_kernel_time: bad argument");
    if (mixed && !p->has_fp32) throw Error(SP_ERR_INVALID, "plain_kernel_time: FP32 values not prepared");
    const DevCsr& M = side == 0 ? p->at : p->a;
    if (!M.nblocks) return 0.0;
    // With column panels a side is several launches (panel sums, short rows,
    // finishing); they are timed together. Otherwise the main kernel alone.
    const bool whole = M.pmat != nullptr;
    // Copies of what the kernel writes: the iterate it updates in place (and
    // x̄ on the A'ᵀ side).
    std::vector<void*> scratch;
    auto copy_of = [&](const void* src, size_t bytes) {
        void* d = dmalloc<char>(std::max<size_t>(bytes, 1));
        scratch.push_back(d);
        if (bytes) hip_check(hipMemcpyAsync(d, src, bytes, hipMemcpyDeviceToDevice, b->stream), "hipMemcpyAsync D2D");
        return d;
    };
    const size_t n = (size_t)p->n, m = (size_t)p->m;
    double ms = 0.0;
    try {
        if (!mixed && side == 0) {
            double* x = (double*)copy_of(a.x->d, 8 * n);
            double* xb = (double*)copy_of(a.x_bar->d, 8 * n);
            auto once = [&] {
                if (whole)
                    spmv_launch(b, b->k_spmv_t_primal_halpern, b->k_fin_t_primal_halpern, M, M.va, a.y->d, x,
                                (const double*)p->c, (const double*)p->lv, (const double*)p->uv,
                                (const double*)a.x_anchor->d, a.tau, a.w, a.rho, xb);
                else
                    spmv_main(b, b->k_spmv_t_primal_halpern, M, (const void*)M.va, (const void*)a.y->d, x,
                              (const double*)p->c, (const double*)p->lv, (const double*)p->uv,
                              (const double*)a.x_anchor->d, a.tau, a.w, a.rho, xb);
            };
            once();
            ms = time_reps(b, reps, once) / reps;
        } else if (!mixed) {
            double* y = (double*)copy_of(a.y->d, 8 * m);
            auto once = [&] {
                if (whole)
                    spmv_launch(b, b->k_spmv_dual_halpern, b->k_fin_dual_halpern, M, M.va, a.x_bar->d, y,
                                (const double*)p->lc, (const double*)p->uc, (const double*)a.y_anchor->d, a.sigma,
                                a.w, a.rho);
                else
                    launch(b, b->k_spmv_dual_halpern, (unsigned)M.nblocks, M.desc, M.rp, M.ci, (const void*)M.va,
                           (const void*)a.x_bar->d, (void*)M.partials, y, (const double*)p->lc, (const double*)p->uc,
                           (const double*)a.y_anchor->d, a.sigma, a.w, a.rho);
            };
            once();
            ms = time_reps(b, reps, once) / reps;
        } else if (side == 0) {
            float* dx = (float*)copy_of(a.dx->d, 4 * n);
            float* dxb = (float*)copy_of(a.dx_bar->d, 4 * n);
            auto once = [&] {
                if (whole)
                    spmv_launch(b, b->k_spmv_t_primal_mixed, b->k_fin_t_primal_mixed, M, M.va32, a.dy->d, dx,
                                (const double*)a.x->d, (const double*)p->c, (const double*)p->lv,
                                (const double*)p->uv, (const double*)a.atya->d, (const float*)a.ex->d, a.tau, a.w,
                                a.rho, dxb);
                else
                    spmv_main(b, b->k_spmv_t_primal_mixed, M, (const void*)M.va32, (const void*)a.dy->d, dx,
                              (const double*)a.x->d, (const double*)p->c, (const double*)p->lv,
                              (const double*)p->uv, (const double*)a.atya->d, (const float*)a.ex->d, a.tau, a.w,
                              a.rho, dxb);
            };
            once();
            ms = time_reps(b, reps, once) / reps;
        } else {
            float* dy = (float*)copy_of(a.dy->d, 4 * m);
            auto once = [&] {
                if (whole)
                    spmv_launch(b, b->k_spmv_dual_mixed, b->k_fin_dual_mixed, M, M.va32, a.dx_bar->d, dy,
                                (const double*)a.y->d, (const double*)p->lc, (const double*)p->uc,
                                (const double*)a.axa->d, (const float*)a.ey->d, a.sigma, a.w, a.rho);
                else
                    launch(b, b->k_spmv_dual_mixed, (unsigned)M.nblocks, M.desc, M.rp, M.ci, (const void*)M.va32,
                           (const void*)a.dx_bar->d, (void*)M.partials, dy, (const double*)a.y->d,
                           (const double*)p->lc, (const double*)p->uc, (const double*)a.axa->d,
                           (const float*)a.ey->d, a.sigma, a.w, a.rho);
            };
            once();
            ms = time_reps(b, reps, once) / reps;
        }
    } catch (...) {
        (void)hipStreamSynchronize(b->stream);
        for (void* q : scratch) dfree(q);
        throw;
    }
    (void)hipStreamSynchronize(b->stream);
    for (void* q : scratch) dfree(q);
    return ms;
}

// ---- SpMV --------------------------------------------------------------------

void Hip::spmv(const sp_problem* p, const sp_vec* x, sp_vec* out) {
    need_len(x, p->n, "spmv x");
    need_len(out, p->m, "spmv out");
    spmv_launch(b, b->k_spmv, b->k_fin, p->a, p->a.va, x->d, out->d);
}

void Hip::spmv_t(const sp_problem* p, const sp_vec* y, sp_vec* out) {
    need_len(y, p->m, "spmv_t y");
    need_len(out, p->n, "spmv_t out");
    spmv_launch(b, b->k_spmv, b->k_fin, p->at, p->at.va, y->d, out->d);
}

void Hip::spmv_t_primal(const sp_problem* p, const sp_vec* y, const sp_vec* x, double tau, sp_vec* aty,
                        sp_vec* x_t, sp_vec* x_bar) {
    need_len(y, p->m, "spmv_t_primal y");
    for (const sp_vec* v : {x, (const sp_vec*)aty, (const sp_vec*)x_t, (const sp_vec*)x_bar})
        need_len(v, p->n, "spmv_t_primal");
    spmv_launch(b, b->k_spmv_t_primal, b->k_fin_t_primal, p->at, p->at.va, y->d, (const double*)x->d,
                (const double*)p->c, (const double*)p->lv, (const double*)p->uv, tau, aty->d, x_t->d, x_bar->d);
}

void Hip::spmv_dual(const sp_problem* p, const sp_vec* x_bar, const sp_vec* y, double sigma, sp_vec* ax,
                    sp_vec* y_t) {
    need_len(x_bar, p->n, "spmv_dual x_bar");
    for (const sp_vec* v : {y, (const sp_vec*)ax, (const sp_vec*)y_t}) need_len(v, p->m, "spmv_dual");
    spmv_launch(b, b->k_spmv_dual, b->k_fin_dual, p->a, p->a.va, x_bar->d, (const double*)y->d,
                (const double*)p->lc, (const double*)p->uc, sigma, ax->d, y_t->d);
}

void Hip::plain_iteration(const sp_problem* p, sp_vec* x, sp_vec* y, const sp_vec* x_anchor,
                          const sp_vec* y_anchor, sp_vec* x_bar, double tau, double sigma, double w, double rho) {
    spmv_launch(b, b->k_spmv_t_primal_halpern, b->k_fin_t_primal_halpern, p->at, p->at.va, y->d, x->d,
                (const double*)p->c, (const double*)p->lv, (const double*)p->uv, (const double*)x_anchor->d, tau, w,
                rho, x_bar->d);
    spmv_launch(b, b->k_spmv_dual_halpern, b->k_fin_dual_halpern, p->a, p->a.va, x_bar->d, y->d,
                (const double*)p->lc, (const double*)p->uc, (const double*)y_anchor->d, sigma, w, rho);
}

// ---- element-wise ------------------------------------------------------------------

void Hip::primal_step(const sp_problem* p, const sp_vec* x, const sp_vec* aty, double tau, sp_vec* x_t,
                      sp_vec* x_bar) {
    for (const sp_vec* v : {x, aty, (const sp_vec*)x_t, (const sp_vec*)x_bar}) need_len(v, p->n, "primal_step");
    if (p->n)
        launch(b, b->k_primal_step, ew_grid(p->n), (const double*)x->d, (const double*)aty->d, (const double*)p->c,
               (const double*)p->lv, (const double*)p->uv, tau, x_t->d, x_bar->d, (long long)p->n);
}

void Hip::dual_step(const sp_problem* p, const sp_vec* y, const sp_vec* ax_bar, double sigma, sp_vec* y_t) {
    for (const sp_vec* v : {y, ax_bar, (const sp_vec*)y_t}) need_len(v, p->m, "dual_step");
    if (p->m)
        launch(b, b->k_dual_step, ew_grid(p->m), (const double*)y->d, (const double*)ax_bar->d, (const double*)p->lc,
               (const double*)p->uc, sigma, y_t->d, (long long)p->m);
}

void Hip::halpern_step(sp_vec* z, const sp_vec* z_t, const sp_vec* anchor, double w, double rho) {
    if (!z) throw Error(SP_ERR_INVALID, "halpern_step: null vector");
    need_len(z_t, z->len, "halpern_step z_t");
    need_len(anchor, z->len, "halpern_step anchor");
    if (z->len)
        launch(b, b->k_halpern, ew_grid(z->len), z->d, (const double*)z_t->d, (const double*)anchor->d, w, rho,
               (long long)z->len);
}

void Hip::sub(const sp_vec* a, const sp_vec* bv, sp_vec* out) {
    if (!a) throw Error(SP_ERR_INVALID, "sub: null vector");
    need_len(bv, a->len, "sub b");
    need_len(out, a->len, "sub out");
    if (a->len)
        launch(b, b->k_sub, ew_grid(a->len), (const double*)a->d, (const double*)bv->d, out->d, (long long)a->len);
}

void Hip::unscale(const sp_vec* v, const double* f, int64_t flen, double s, const int* to, bool dual, sp_vec* out) {
    if (!v || (v->len && !f)) throw Error(SP_ERR_INVALID, "unscale: null argument");
    if (flen != v->len) throw Error(SP_ERR_SHAPE, "unscale: factors of the wrong length");
    need_len(out, v->len, "unscale out");
    if (v->len)
        launch(b, dual ? b->k_unscale_dual : b->k_unscale_primal, ew_grid(v->len), (const double*)v->d, f, s, to,
               out->d, (long long)v->len);
}

void Hip::unscale_to_host(const sp_vec* v, const std::vector<double>& f, double s, const std::vector<int64_t>* order,
                          bool dual, double* out) {
    if (!v) throw Error(SP_ERR_INVALID, "unscale: null vector");
    const int64_t n = v->len;
    if ((int64_t)f.size() != n) throw Error(SP_ERR_SHAPE, "unscale: factors of the wrong length");
    const bool relabeled = order && !order->empty();
    if (relabeled && (int64_t)order->size() != n) throw Error(SP_ERR_SHAPE, "unscale: relabeling of the wrong length");
    if (n == 0) return;
    std::vector<int> to32;
    if (relabeled) {
        to32.resize((size_t)n);
        for (int64_t q = 0; q < n; ++q) {
            const int64_t t = (*order)[(size_t)q];
            if (t < 0 || t >= n) throw Error(SP_ERR_INVALID, "unscale: relabeling out of range");
            to32[(size_t)q] = (int)t;
        }
    }
    double* fd = nullptr;
    int* td = nullptr;
    sp_vec* o = nullptr;
    try {
        fd = dupload(b->stream, f.data(), f.size());
        if (relabeled) td = dupload(b->stream, to32.data(), to32.size());
        o = zeros(n);
        unscale(v, fd, n, s, td, dual, o);
        download(o, out, n);
    } catch (...) {
        (void)hipStreamSynchronize(b->stream);
        dfree(fd);
        dfree(td);
        free_vec(o);
        throw;
    }
    dfree(fd);
    dfree(td);
    free_vec(o);
}

void Hip::scale_inv_sqrt(sp_vec* v, const double* ns) {
    if (!v || !ns) throw Error(SP_ERR_INVALID, "scale_inv_sqrt: null argument");
    if (v->len) launch(b, b->k_scale_inv_sqrt, ew_grid(v->len), v->d, ns, (long long)v->len);
}

void Hip::scale(sp_vec* v, double alpha) {
    if (!v) throw Error(SP_ERR_INVALID, "scale: null vector");
    if (v->len) launch(b, b->k_scale, ew_grid(v->len), v->d, alpha, (long long)v->len);
}

// ---- reductions ------------------------------------------------------------------------

int Hip::reduce_grid(int64_t n) const {
    // A fixed function of n, so a given call always reduces in the same order.
    int64_t g = (n + 8 * kWG - 1) / (8 * kWG);
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, kMaxReduceBlocks));
}

// Launches a reduction over `n` elements whose K totals land in `dst`
// (device memory): per-workgroup partials, then the finishing pass.
template <class... Args>
void reduce_into(sp_backend* b, hipFunction_t k, int grid, int K, double* dst, Args... args) {
    launch(b, k, (unsigned)grid, args..., b->red_partials, dst);
    if (grid > 1) launch(b, b->k_red_finish, 1u, (const double*)b->red_partials, grid, K, dst);
}

const double* Hip::read_result(int k, int grid) {
    (void)grid;
    hip_check(hipMemcpyAsync(b->host_result, b->red_result, sizeof(double) * (size_t)k, hipMemcpyDeviceToHost,
                             b->stream),
              "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    return b->host_result;
}

double* Hip::slot(int i) {
    if (i < 0 || i >= kSlots) throw Error(SP_ERR_INVALID, "result slot out of range");
    return b->slots + i;
}

const double* Hip::fetch(int count) {
    if (count < 0 || count > kSlots) throw Error(SP_ERR_INVALID, "fetch: bad slot count");
    hip_check(hipMemcpyAsync(b->host_slots, b->slots, sizeof(double) * (size_t)count, hipMemcpyDeviceToHost,
                             b->stream),
              "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    return b->host_slots;
}

const double* Hip::fetch_vec(const sp_vec* v, int count) {
    if (!v || count < 0 || count > kSlots || count > v->len) throw Error(SP_ERR_INVALID, "fetch: bad slot count");
    hip_check(hipMemcpyAsync(b->host_slots, v->d, sizeof(double) * (size_t)count, hipMemcpyDeviceToHost, b->stream),
              "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    return b->host_slots;
}

void Hip::fetch_vec_async(const sp_vec* v, int count, double* dst) {
    if (!v || !dst || count < 0 || count > v->len) throw Error(SP_ERR_INVALID, "fetch: bad slot count");
    hip_check(hipMemcpyAsync(dst, v->d, sizeof(double) * (size_t)count, hipMemcpyDeviceToHost, b->stream),
              "hipMemcpyAsync D2H");
}

double* Hip::host_alloc(int64_t count) {
    void* p = nullptr;
    hip_check(hipHostMalloc(&p, sizeof(double) * (size_t)std::max<int64_t>(count, 1), hipHostMallocDefault),
              "hipHostMalloc");
    return static_cast<double*>(p);
}

void free_host(double* h) {
    if (h) (void)hipHostFree(h);
}

// Raw sums -> the trait's structs (cpu/src/lib.rs:211-259).
sp_primal_stats_t Hip::primal_from(const double* r) { return sp_primal_stats_t{r[0], r[1]}; }
// dual_stats: objective = rows + cols (cpu/src/lib.rs:223).
sp_dual_stats_t Hip::dual_from(const double* r) { return sp_dual_stats_t{r[1], r[0] + r[2]}; }
// primal_ray_stats: violation = rows + cols (cpu/src/lib.rs:245).
sp_ray_stats_t Hip::primal_ray_from(const double* r) { return sp_ray_stats_t{r[3], r[0] + r[1], r[2]}; }
// dual_ray_stats: objective = row_obj + col_obj, violation = row_viol + col_viol (cpu/src/lib.rs:258).
sp_ray_stats_t Hip::dual_ray_from(const double* r) { return sp_ray_stats_t{r[1] + r[4], r[0] + r[3], r[2]}; }

void Hip::dot_into(const sp_vec* a, const sp_vec* bv, double* dst) {
    if (!a) throw Error(SP_ERR_INVALID, "dot: null vector");
    need_len(bv, a->len, "dot");
    reduce_into(b, b->k_dot, reduce_grid(a->len), 1, dst, (const double*)a->d, (const double*)bv->d,
                (long long)a->len);
}

void Hip::distance_sq_into(const sp_vec* a, const sp_vec* bv, double* dst) {
    if (!a) throw Error(SP_ERR_INVALID, "distance_sq: null vector");
    need_len(bv, a->len, "distance_sq");
    reduce_into(b, b->k_dist, reduce_grid(a->len), 1, dst, (const double*)a->d, (const double*)bv->d,
                (long long)a->len);
}

void Hip::primal_stats_into(const sp_problem* p, const sp_vec* x, const sp_vec* ax, double* dst) {
    need_len(x, p->n, "primal_stats x");
    need_len(ax, p->m, "primal_stats ax");
    reduce_into(b, b->k_primal_stats, reduce_grid(p->m + p->n), 2, dst, (const double*)ax->d, (const double*)p->lc,
                (const double*)p->uc, (const double*)p->pf, (long long)p->m, (const double*)p->c,
                (const double*)x->d, (long long)p->n);
}

void Hip::dual_stats_into(const sp_problem* p, const sp_vec* y, const sp_vec* aty, double* dst) {
    need_len(y, p->m, "dual_stats y");
    need_len(aty, p->n, "dual_stats aty");
    reduce_into(b, b->k_dual_stats, reduce_grid(p->m + p->n), 3, dst, (const double*)y->d, (const double*)p->lc,
                (const double*)p->uc, (long long)p->m, (const double*)p->c, (const double*)aty->d,
                (const double*)p->lv, (const double*)p->uv, (const double*)p->df, (long long)p->n);
}

void Hip::step_dual_residual_sq_into(const sp_problem* p, const sp_vec* x, const sp_vec* x_t, const sp_vec* aty,
                                     const sp_vec* aty_t, double tau, double* dst) {
    for (const sp_vec* v : {x, x_t, aty, aty_t}) need_len(v, p->n, "step_dual_residual_sq");
    const double inv_tau = 1.0 / tau;  // cpu/src/lib.rs:235: multiply, don't divide
    reduce_into(b, b->k_step_dual, reduce_grid(p->n), 1, dst, (const double*)x->d, (const double*)x_t->d,
                (const double*)aty->d, (const double*)aty_t->d, (const double*)p->df, inv_tau, (long long)p->n);
}

void Hip::primal_ray_stats_into(const sp_problem* p, const sp_vec* dx, const sp_vec* adx, double* dst) {
    need_len(dx, p->n, "primal_ray_stats dx");
    need_len(adx, p->m, "primal_ray_stats adx");
    reduce_into(b, b->k_primal_ray, reduce_grid(p->m + p->n), 4, dst, (const double*)adx->d, (const double*)p->lc,
                (const double*)p->uc, (long long)p->m, (const double*)dx->d, (const double*)p->c,
                (const double*)p->lv, (const double*)p->uv, (long long)p->n);
}

void Hip::dual_ray_stats_into(const sp_problem* p, const sp_vec* dy, const sp_vec* atdy, double* dst) {
    need_len(dy, p->m, "dual_ray_stats dy");
    need_len(atdy, p->n, "dual_ray_stats atdy");
    reduce_into(b, b->k_dual_ray, reduce_grid(p->m + p->n), 5, dst, (const double*)dy->d, (const double*)p->lc,
                (const double*)p->uc, (long long)p->m, (const double*)atdy->d, (const double*)p->lv,
                (const double*)p->uv, (long long)p->n);
}

// ---- an OpList on the device --------------------------------------------------------

struct DevOpList {
    // Host copies of the tables (alive while their asynchronous uploads may
    // still read them) and the device tables.
    std::vector<COp> h_ops;
    std::vector<CBlk> h_sub, h_red, h_hal;
    std::vector<int> h_fin;
    COp* ops = nullptr;
    CBlk *sub = nullptr, *red = nullptr, *hal = nullptr;
    int* fin = nullptr;
    double* partials = nullptr;
    // step_dual entries: their 1/τ (s0) is rewritten when τ changes, from a
    // ring of pinned staging slots; a slot is reused once the event recorded
    // after its copy has completed.
    static constexpr int kTauSlots = 8;
    std::vector<int> step_ops;
    double inv_tau = 0.0;
    double* h_tau = nullptr;
    hipEvent_t tau_done[kTauSlots] = {};
    bool tau_used[kTauSlots] = {};
    int tau_next = 0;
};

int red_sums(int kind) {
    switch (kind) {
        case kRedDot:
        case kRedDist:
        case kRedStep:
            return 1;
        case kRedPrimal:
            return 2;
        case kRedDual:
            return 3;
        case kRedPrimalRay:
            return 4;
        case kRedDualRay:
            return 5;
    }
    throw Error(SP_ERR_INVALID, "unknown reduction kind");
}

DevOpList* Hip::upload_ops(const OpList& l) {
    auto d = std::make_unique<DevOpList>();
    auto tiles = [](std::vector<CBlk>& blk, const COp& o, int i) {
        for (long long t = 0; t * kCTile < o.n; ++t) blk.push_back(CBlk{i, (int)t});
    };
    for (const COp& o : l.subs) {
        tiles(d->h_sub, o, (int)d->h_ops.size());
        d->h_ops.push_back(o);
    }
    int pb = 0;
    for (COp o : l.reds) {
        // The workgroup count of the method's own launch (reduce_into).
        o.K = red_sums(o.kind);
        o.G = reduce_grid(red_len(o));
        o.pbase = pb;
        pb += o.G * o.K;
        const int i = (int)d->h_ops.size();
        for (int t = 0; t < o.G; ++t) d->h_red.push_back(CBlk{i, t});
        if (o.G > 1) d->h_fin.push_back(i);
        if (o.kind == kRedStep) {
            d->step_ops.push_back(i);
            d->inv_tau = o.s0;
        }
        d->h_ops.push_back(o);
    }
    for (const COp& o : l.halperns) {
        tiles(d->h_hal, o, (int)d->h_ops.size());
        d->h_ops.push_back(o);
    }
    hipStream_t st = b->stream;
    try {
        d->ops = dupload(st, d->h_ops.data(), d->h_ops.size());
        d->sub = dupload(st, d->h_sub.data(), d->h_sub.size());
        d->red = dupload(st, d->h_red.data(), d->h_red.size());
        d->hal = dupload(st, d->h_hal.data(), d->h_hal.size());
        d->fin = dupload(st, d->h_fin.data(), d->h_fin.size());
        d->partials = dmalloc<double>((size_t)std::max(pb, 1));
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&d->h_tau), sizeof(double) * DevOpList::kTauSlots,
                                hipHostMallocDefault),
                  "hipHostMalloc");
        for (hipEvent_t& e : d->tau_done) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    } catch (...) {
        free_ops(d.release());
        throw;
    }
    return d.release();
}

void free_ops(DevOpList* d) {
    if (!d) return;
    dfree(d->ops);
    dfree(d->sub);
    dfree(d->red);
    dfree(d->hal);
    dfree(d->fin);
    dfree(d->partials);
    if (d->h_tau) (void)hipHostFree(d->h_tau);
    for (hipEvent_t e : d->tau_done)
        if (e) (void)hipEventDestroy(e);
    delete d;
}

void Hip::run_subs(const DevOpList* d) {
    if (!d->h_sub.empty()) launch(b, b->k_csub, (unsigned)d->h_sub.size(), (const CBlk*)d->sub, (const COp*)d->ops);
}

void Hip::run_reductions(DevOpList* d, double inv_tau) {
    if (d->h_red.empty()) return;
    if (!d->step_ops.empty() && !(inv_tau == d->inv_tau)) {
        const int k = d->tau_next++ % DevOpList::kTauSlots;
        if (d->tau_used[k]) hip_check(hipEventSynchronize(d->tau_done[k]), "hipEventSynchronize");  // its last copy ran
        d->h_tau[k] = inv_tau;
        for (int i : d->step_ops)
            hip_check(hipMemcpyAsync(reinterpret_cast<char*>(d->ops + i) + offsetof(COp, s0), d->h_tau + k,
                                     sizeof(double), hipMemcpyHostToDevice, b->stream),
                      "hipMemcpyAsync H2D");
        hip_check(hipEventRecord(d->tau_done[k], b->stream), "hipEventRecord");
        d->tau_used[k] = true;
        d->inv_tau = inv_tau;
    }
    launch(b, b->k_cred_any, (unsigned)d->h_red.size(), (const CBlk*)d->red, (const COp*)d->ops, d->partials);
    if (!d->h_fin.empty())
        launch(b, b->k_cred_finish_any, (unsigned)d->h_fin.size(), (const int*)d->fin, (const COp*)d->ops,
               (const double*)d->partials);
}

void Hip::run_halperns(const DevOpList* d, double w, double rho) {
    if (!d->h_hal.empty())
        launch(b, b->k_chalpern_wr, (unsigned)d->h_hal.size(), (const CBlk*)d->hal, (const COp*)d->ops, w, rho);
}

double Hip::dot(const sp_vec* a, const sp_vec* bv) {
    dot_into(a, bv, b->red_result);
    return read_result(1, 0)[0];
}

double Hip::norm_sq(const sp_vec* v) { return dot(v, v); }

double Hip::distance_sq(const sp_vec* a, const sp_vec* bv) {
    distance_sq_into(a, bv, b->red_result);
    return read_result(1, 0)[0];
}

sp_primal_stats_t Hip::primal_stats(const sp_problem* p, const sp_vec* x, const sp_vec* ax) {
    primal_stats_into(p, x, ax, b->red_result);
    return primal_from(read_result(2, 0));
}

sp_dual_stats_t Hip::dual_stats(const sp_problem* p, const sp_vec* y, const sp_vec* aty) {
    dual_stats_into(p, y, aty, b->red_result);
    return dual_from(read_result(3, 0));
}

sp_dual_stats_t Hip::dual_stats_in(const sp_problem* p, const sp_vec* y, const sp_vec* aty, double* ws,
                                   int64_t cap) {
    need_len(y, p->m, "dual_stats y");
    need_len(aty, p->n, "dual_stats aty");
    constexpr int K = 3;
    const int grid = reduce_grid(p->m + p->n);
    // Partials first, the K totals in the last 8 doubles.
    if (!ws || (int64_t)grid * K > cap - 8)
        throw Error(SP_ERR_TOO_LARGE, "dual_stats: the reduction needs " + std::to_string((int64_t)grid * K + 8) +
                                          " doubles of scratch, more than the " + std::to_string(cap) + " it has");
    double* partials = ws;
    double* result = ws + cap - 8;
    auto go = [&](hipFunction_t f, unsigned g, auto... args) {
        void* params[] = {static_cast<void*>(&args)...};
        hip_check(hipModuleLaunchKernel(f, g, 1, 1, kWG, 1, 1, 0, b->stream, params, nullptr),
                  "hipModuleLaunchKernel");
    };
    go(b->k_dual_stats, (unsigned)grid, (const double*)y->d, (const double*)p->lc, (const double*)p->uc,
       (long long)p->m, (const double*)p->c, (const double*)aty->d, (const double*)p->lv, (const double*)p->uv,
       (const double*)p->df, (long long)p->n, partials, result);
    if (grid > 1) go(b->k_red_finish, 1u, (const double*)partials, grid, K, result);
    double r[K];
    hip_check(hipMemcpyAsync(r, result, sizeof r, hipMemcpyDeviceToHost, b->stream), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    return dual_from(r);
}

double Hip::step_dual_residual_sq(const sp_problem* p, const sp_vec* x, const sp_vec* x_t, const sp_vec* aty,
                                  const sp_vec* aty_t, double tau) {
    step_dual_residual_sq_into(p, x, x_t, aty, aty_t, tau, b->red_result);
    return read_result(1, 0)[0];
}

sp_ray_stats_t Hip::primal_ray_stats(const sp_problem* p, const sp_vec* dx, const sp_vec* adx) {
    primal_ray_stats_into(p, dx, adx, b->red_result);
    return primal_ray_from(read_result(4, 0));
}

sp_ray_stats_t Hip::dual_ray_stats(const sp_problem* p, const sp_vec* dy, const sp_vec* atdy) {
    dual_ray_stats_into(p, dy, atdy, b->red_result);
    return dual_ray_from(read_result(5, 0));
}

}  // namespace sp

// ---- lifecycle (outside the namespace: used by the C ABI) ----------------------------

namespace {

std::string code_object_path() {
    if (const char* env = std::getenv("SADDLEPOINT_HIP_CODE_OBJECT")) return env;
    Dl_info info{};
    if (dladdr(reinterpret_cast<void*>(&sp::hip_check), &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t slash = p.rfind('/');
        return (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/saddlepoint_kernels.hsaco";
    }
    return "saddlepoint_kernels.hsaco";
}

}  // namespace

sp_backend* sp_backend_open(int ordinal) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        throw sp::Error(SP_ERR_UNAVAILABLE, "HIP unavailable: no GPU device");
    if (ordinal < 0 || ordinal >= count)
        throw sp::Error(SP_ERR_UNAVAILABLE, "HIP unavailable: no device " + std::to_string(ordinal));
    auto* b = new sp_backend;
    b->serial = sp::next_backend_serial();
    b->device = ordinal;
    try {
        sp::hip_check(hipSetDevice(ordinal), "hipSetDevice");
        hipDeviceProp_t prop{};
        sp::hip_check(hipGetDeviceProperties(&prop, ordinal), "hipGetDeviceProperties");
        b->cus = prop.multiProcessorCount;
        const std::string dev = prop.name[0] ? std::string(prop.name) : "device " + std::to_string(ordinal);
        b->name = "hip (" + dev + ", " + prop.gcnArchName + ")";
        // One default-priority stream carries every operation of this backend.
        sp::hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        const std::string path = code_object_path();
        if (hipModuleLoad(&b->module, path.c_str()) != hipSuccess)
            throw sp::Error(SP_ERR_UNAVAILABLE, "HIP unavailable: cannot load code object " + path);
        auto get = [&](hipFunction_t& f, const char* name) {
            sp::hip_check(hipModuleGetFunction(&f, b->module, name), name);
            b->kernel_fns.push_back(f);
            b->kernel_names.push_back(name);
            b->kernel_ms.push_back(0.0);
            b->kernel_launches.push_back(0);
        };
        get(b->k_spmv, "sp_spmv_f64");
        get(b->k_fin, "sp_finish_f64");
        get(b->k_fin_t_primal, "sp_finish_t_primal_f64");
        get(b->k_fin_t_primal_halpern, "sp_finish_t_primal_halpern_f64");
        get(b->k_fin_dual, "sp_finish_dual_f64");
        get(b->k_fin_dual_halpern, "sp_finish_dual_halpern_f64");
        get(b->k_fin32, "sp_finish_f32");
        get(b->k_fin_t_primal_mixed, "sp_finish_t_primal_mixed");
        get(b->k_fin_dual_mixed, "sp_finish_dual_mixed");
        get(b->k_red_finish, "sp_red_finish");
        get(b->k_bspmv_t64, "sp_bspmv_t_primal_halpern_f64");
        get(b->k_bspmv_a64, "sp_bspmv_dual_halpern_f64");
        get(b->k_bfin_t64, "sp_bfinish_t_primal_halpern_f64");
        get(b->k_bfin_a64, "sp_bfinish_dual_halpern_f64");
        get(b->k_bspmv_t_mixed, "sp_bspmv_t_primal_mixed");
        get(b->k_bspmv_a_mixed, "sp_bspmv_dual_mixed");
        get(b->k_bfin_t_mixed, "sp_bfinish_t_primal_mixed");
        get(b->k_bfin_a_mixed, "sp_bfinish_dual_mixed");
        get(b->k_cspmv_store, "sp_cspmv_store_f64");
        get(b->k_cspmv_primal, "sp_cspmv_primal_f64");
        get(b->k_cspmv_dual, "sp_cspmv_dual_f64");
        get(b->k_cfin_store, "sp_cfinish_store_f64");
        get(b->k_cfin_primal, "sp_cfinish_primal_f64");
        get(b->k_cfin_dual, "sp_cfinish_dual_f64");
        get(b->k_csub, "sp_csub_f64");
        get(b->k_chalpern, "sp_chalpern_f64");
        get(b->k_cdelta_to, "sp_cdelta_to");
        get(b->k_cdelta_from, "sp_cdelta_from");
        get(b->k_ccopy, "sp_ccopy_f64");
        get(b->k_cred_any, "sp_cred_any_f64");
        get(b->k_cred_finish_any, "sp_cred_finish_any");
        get(b->k_chalpern_wr, "sp_chalpern_wr_f64");
        get(b->k_cscale_inv_sqrt, "sp_cscale_inv_sqrt_f64");
        get(b->k_sc_minmax, "sp_sc_minmax");
        get(b->k_sc_powsum, "sp_sc_powsum");
        get(b->k_sc_factor, "sp_sc_factor");
        get(b->k_sc_apply, "sp_sc_apply");
        get(b->k_sc_mul, "sp_sc_mul");
        get(b->k_tr_rows, "sp_tr_rows");
        get(b->k_tr_hist, "sp_tr_hist");
        get(b->k_scan_tiles, "sp_scan_tiles");
        get(b->k_scan_add, "sp_scan_add");
        get(b->k_tr_scatter, "sp_tr_scatter");
        get(b->k_tr_ptr, "sp_tr_ptr");
        get(b->k_tr_gather, "sp_tr_gather");
        get(b->k_tr_iota, "sp_tr_iota");
        get(b->k_remap, "sp_remap");
        get(b->k_col_count, "sp_col_count");
        get(b->k_scale_inv_sqrt, "sp_scale_inv_sqrt_f64");
        get(b->k_red_check, "sp_red_check_f64");
        get(b->k_red_finish7, "sp_red_finish7");
        get(b->k_spmv_t_primal, "sp_spmv_t_primal_f64");
        get(b->k_spmv_t_primal_halpern, "sp_spmv_t_primal_halpern_f64");
        get(b->k_spmv_dual, "sp_spmv_dual_f64");
        get(b->k_spmv_dual_halpern, "sp_spmv_dual_halpern_f64");
        get(b->k_primal_step, "sp_primal_step_f64");
        get(b->k_dual_step, "sp_dual_step_f64");
        get(b->k_halpern, "sp_halpern_step_f64");
        get(b->k_sub, "sp_sub_f64");
        get(b->k_scale, "sp_scale_f64");
        get(b->k_fill, "sp_fill_f64");
        get(b->k_unscale_primal, "sp_unscale_primal_f64");
        get(b->k_unscale_dual, "sp_unscale_dual_f64");
        get(b->k_spmv32, "sp_spmv_f32");
        get(b->k_probe_copy, "sp_probe_copy");
        get(b->k_probe_triad, "sp_probe_triad");
        get(b->k_probe_gather, "sp_probe_gather");
        get(b->k_probe_copy32, "sp_probe_copy_f32");
        get(b->k_probe_triad32, "sp_probe_triad_f32");
        get(b->k_probe_overlap, "sp_probe_overlap");
        get(b->k_replay64, "sp_probe_replay_f64");
        get(b->k_panel_count, "sp_panel_count");
        get(b->k_panel_gather, "sp_panel_gather");
        get(b->k_spmv_panel, "sp_spmv_panel_f64");
        get(b->k_fin_panel, "sp_finish_panel_f64");
        get(b->k_spmv_panel32, "sp_spmv_panel_f32");
        get(b->k_fin_panel32, "sp_finish_panel_f32");
        get(b->k_rows_gather, "sp_rows_gather");
        get(b->k_gather, "sp_gather_f64");
        get(b->k_replay32, "sp_probe_replay_f32");
        get(b->k_spmv_t_primal_mixed, "sp_spmv_t_primal_mixed");
        get(b->k_spmv_dual_mixed, "sp_spmv_dual_mixed");
        get(b->k_spmv_t_primal_halpern_hot, "sp_spmv_t_primal_halpern_hot_f64");
        get(b->k_spmv_t_primal_mixed_hot, "sp_spmv_t_primal_mixed_hot");
        // Their persistent grids: workgroups per CU (0, if the runtime cannot
        // say, keeps the plain kernels).
        if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&b->hot_occ64, b->k_spmv_t_primal_halpern_hot, sp::kWG,
                                                               0) != hipSuccess)
            b->hot_occ64 = 0;
        if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&b->hot_occ32, b->k_spmv_t_primal_mixed_hot, sp::kWG,
                                                               0) != hipSuccess)
            b->hot_occ32 = 0;
        get(b->k_to_f32, "sp_f64_to_f32");
        get(b->k_delta_from, "sp_delta_from");
        get(b->k_delta_to, "sp_delta_to");
        get(b->k_dot, "sp_red_dot_f64");
        get(b->k_dist, "sp_red_distance_sq_f64");
        get(b->k_primal_stats, "sp_red_primal_stats_f64");
        get(b->k_dual_stats, "sp_red_dual_stats_f64");
        get(b->k_step_dual, "sp_red_step_dual_f64");
        get(b->k_primal_ray, "sp_red_primal_ray_f64");
        get(b->k_dual_ray, "sp_red_dual_ray_f64");
        // The no-allocation entry point's scratch: this device's copy of
        // the static array, its address taken now with the device current.
        try {
            b->scratch = sp::scratch_on_device(ordinal);
        } catch (const sp::Error&) {
            b->scratch = nullptr;  // sp_dual_stats_noalloc then reports SP_ERR_UNAVAILABLE
        }
        b->red_partials = sp::dmalloc<double>((size_t)sp::Hip::kMaxReduceBlocks * sp::Hip::kMaxSlots);
        b->red_result = sp::dmalloc<double>(sp::Hip::kMaxSlots);
        b->slots = sp::dmalloc<double>(sp::Hip::kSlots);
        sp::hip_check(hipHostMalloc(reinterpret_cast<void**>(&b->host_slots), sizeof(double) * sp::Hip::kSlots,
                                    hipHostMallocDefault),
                      "hipHostMalloc");
        sp::hip_check(hipHostMalloc(reinterpret_cast<void**>(&b->host_result), sizeof(double) * sp::Hip::kMaxSlots,
                                    hipHostMallocDefault),
                      "hipHostMalloc");
        sp::hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    } catch (...) {
        sp_backend_close(b);
        throw;
    }
    return b;
}

void sp_backend_close(sp_backend* b) {
    if (!b) return;
    sp::forget_errors(b);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (const auto& pe : b->pending) b->event_pool.push_back(pe.start), b->event_pool.push_back(pe.stop);
    for (hipEvent_t e : b->event_pool) (void)hipEventDestroy(e);
    sp::dfree(b->red_partials);
    sp::dfree(b->red_result);
    if (b->host_result) (void)hipHostFree(b->host_result);
    if (b->host_slots) (void)hipHostFree(b->host_slots);
    sp::dfree(b->slots);
    if (b->module) (void)hipModuleUnload(b->module);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}
