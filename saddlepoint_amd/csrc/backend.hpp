// The HIP backend: device state behind the opaque C handles, and the C++
// methods the PDHG driver (solver.hpp) calls, one per Backend method of the
// reference (crates/saddlepoint-core/src/backend.rs:43-137).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "host.hpp"
#include "saddlepoint_hip.h"

namespace sp {

// Launch descriptor of one SpMV work item; mirrors `struct Desc` in kernels.hip.
struct Desc {
    int r0, r1, k0, k1;
};

// A row split into several chunks: the finishing kernel adds its chunk sums
// partials[pbase + q·stride], q < nch (stride 0 reads as 1). Mirrors
// `struct LongRow` in kernels.hip.
struct LongRow {
    int row, pbase, nch, stride;
};

// How an SpMV is cut into work items (see plan_spmv). The autotuner searches
// these knobs per matrix.
struct SpmvTuning {
    int block_nnz = 2048;  // nonzeros staged per stream workgroup (<= kernel LDS capacity)
    int long_row = 2048;   // rows longer than this get workgroups of their own
    int max_rows = 1024;   // rows per stream workgroup
};

// A CSR matrix on the device with int32 indices and its SpMV plan.
struct DevCsr {
    int nrows = 0, ncols = 0;
    int64_t nnz = 0;
    int* rp = nullptr;
    int* ci = nullptr;
    double* va = nullptr;
    float* va32 = nullptr;  // FP32 copy of the values, made on first mixed-precision use
    Desc* desc = nullptr;
    int nblocks = 0;
    double* partials = nullptr;  // one per chunk of a multi-chunk row
    int npartials = 0;
    LongRow* longs = nullptr;    // the multi-chunk rows
    int nlong = 0;
    std::vector<Desc> host_desc;       // the plan, kept for batch tables
    std::vector<LongRow> host_longs;
    SpmvTuning tuning;
    // Column panels (Hip::set_panels): the long rows (more than kPanelMinRow
    // nonzeros) are also kept as a panel matrix whose row q = p·L + l holds
    // long row l's nonzeros in column panel p (columns [p·2^pshift,
    // (p+1)·2^pshift)), stored and dispatched panel by panel; its sums land in
    // psum[q] and the finishing table `plongs` ({row, l, P, stride L}) adds
    // each long row's P sums in panel order. The short rows run as the
    // stream blocks `pdesc`.
    DevCsr* pmat = nullptr;
    int plong = 0, npanels = 0, pshift = 0;
    double* psum = nullptr;     // L·P panel sums (floats in mixed precision)
    LongRow* plongs = nullptr;  // {row, l, P, L}
    Desc* pdesc = nullptr;      // the short rows' stream blocks
    int npdesc = 0;
    // Hot tables (Hip::plan_hot): the plain-iteration kernel of this matrix
    // keeps x[0, kHotF64) in LDS (FP64), or x[0, kHotF32) (mixed precision).
    bool hot64 = false, hot32 = false;
};

// Hot-table sizes (entries): SP_HOT_F64 and SP_HOT_F32 in kernels.hip.
constexpr int kHotF64 = 2048;
constexpr int kHotF32 = 4096;
// Hot tables are used (Hip::plan_hot) on matrices of at least kHotMinNnz
// nonzeros (hot_table = -1) with at least kHotMinFrac of them gathering an
// entry of the table.
constexpr int64_t kHotMinNnz = int64_t(1) << 23;
constexpr double kHotMinFrac = 0.5;

// Rows longer than this are gathered by column panel when panels are on.
constexpr int kPanelMinRow = 64;

// Host mirror of `struct COp` in kernels.hip (same layout): one entry of a
// batched operation's table (the operands of an SpMV, element-wise pass or
// reduction; which fields an operation reads is listed at its kernel).
struct COp {
    const int *rp = nullptr, *ci = nullptr;
    const double* va = nullptr;
    double* part = nullptr;
    const double* in = nullptr;
    double *o0 = nullptr, *o1 = nullptr, *o2 = nullptr;
    const double *a0 = nullptr, *a1 = nullptr, *a2 = nullptr, *a3 = nullptr, *a4 = nullptr, *a5 = nullptr,
                 *a6 = nullptr, *a7 = nullptr;
    float* f0 = nullptr;
    double s0 = 0, s1 = 0;
    long long m = 0, n = 0;
    double* dst = nullptr;
    int G = 0, pbase = 0;
    int kind = 0, K = 0;
};
static_assert(sizeof(COp) == 192, "COp layout");
// One workgroup of a batched element-wise pass or reduction (mirrors `struct CBlk`).
struct CBlk {
    int op, b;
};
constexpr int kCTile = 4 * 256;  // CTILE: elements per workgroup of an element-wise pass

// Reduction kinds of sp_cred_any_f64 (RK_* in kernels.hip).
enum RedKind { kRedDot = 0, kRedDist = 1, kRedPrimal = 2, kRedDual = 3, kRedStep = 4, kRedPrimalRay = 5,
               kRedDualRay = 6 };
// A kind's sum count K, and the length of an entry's reduction: rows and
// columns (m + n) for the stats and rays, n for the others.
int red_sums(int kind);
inline int64_t red_len(const COp& o) {
    return (o.kind == kRedPrimal || o.kind == kRedDual || o.kind == kRedPrimalRay || o.kind == kRedDualRay)
               ? o.m + o.n
               : o.n;
}

// A fixed list of element-wise differences (o0 = a0 - a1), reductions and
// Halpern updates (o0 = w((1+ρ)a0 - ρ·o0) + (1-w)a1) over device memory that
// stays where it is: one solver's vectors and result slots. Uploaded once
// (Hip::upload_ops); then each group runs as one launch (the reductions as
// two), and every entry computes what its single-method launch computes, in
// the same order, so the results are bitwise those of the methods.
struct OpList {
    std::vector<COp> subs, reds, halperns;  // reds: kind, operands, m, n, dst
};
struct DevOpList;  // the uploaded tables

// Host-side planner: cuts rows into stream blocks and long-row chunks.
struct SpmvPlan {
    std::vector<Desc> desc;
    std::vector<LongRow> longs;
    int npartials = 0;
};
SpmvPlan plan_spmv(const int64_t* row_ptr, int64_t nrows, const SpmvTuning& t);
// The autotuned rule: a plan chosen from the matrix's row-length features by
// the table that the tuning sweep generated (autotune.cpp).
SpmvTuning autotuned_plan(const int64_t* row_ptr, int64_t nrows);

}  // namespace sp

struct sp_vec {
    sp_backend* b = nullptr;
    double* d = nullptr;
    int64_t len = 0;
};

struct sp_problem {
    sp_backend* b = nullptr;
    int64_t m = 0, n = 0;
    sp::DevCsr a, at;
    double *c = nullptr, *lc = nullptr, *uc = nullptr, *lv = nullptr, *uv = nullptr;
    double *pf = nullptr, *df = nullptr;
    bool has_fp32 = false;  // a.va32 / at.va32 are made on first mixed-precision use
};

// An FP32 device vector (driver-internal: the mixed-precision shadows).
struct sp_vec32 {
    float* d = nullptr;
    int64_t len = 0;
};

struct sp_backend {
    uint64_t serial = 0;  // distinct for every backend opened in this process
    int device = 0;
    hipStream_t stream = nullptr;
    hipModule_t module = nullptr;
    int cus = 0;
    std::string name;
    // Reduction scratch: per-workgroup partials, the result slots and their
    // pinned host mirror.
    double* red_partials = nullptr;
    double* red_result = nullptr;
    double* host_result = nullptr;
    // Result slots for reductions whose values are read back together
    // (one transfer per check iteration in the fused driver).
    double* slots = nullptr;
    double* host_slots = nullptr;
    // This device's copy of the no-allocation entry point's static scratch
    // array (noalloc.hip); null if its address could not be taken.
    double* scratch = nullptr;

    // Kernel timing (sp_backend_profile): while enabled, every launch is
    // bracketed by two events on the backend's stream.
    bool profiling = false;
    struct Pending {
        int kernel;
        hipEvent_t start, stop;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::string> kernel_names;   // index = kernel id
    std::vector<hipFunction_t> kernel_fns;
    std::vector<double> kernel_ms;
    std::vector<int64_t> kernel_launches;

    hipFunction_t k_fin, k_fin_t_primal, k_fin_t_primal_halpern, k_fin_dual, k_fin_dual_halpern, k_fin32,
        k_fin_t_primal_mixed, k_fin_dual_mixed, k_red_finish;
    hipFunction_t k_bspmv_t64, k_bspmv_a64, k_bfin_t64, k_bfin_a64, k_bspmv_t_mixed, k_bspmv_a_mixed,
        k_bfin_t_mixed, k_bfin_a_mixed;
    // Batched check operations (serving loop).
    hipFunction_t k_cspmv_store, k_cspmv_primal, k_cspmv_dual, k_cfin_store, k_cfin_primal, k_cfin_dual, k_csub,
        k_chalpern, k_cdelta_to, k_cdelta_from, k_ccopy, k_cscale_inv_sqrt;
    // Reductions of several kinds in one launch (serving phases, Hip::upload_ops).
    hipFunction_t k_cred_any, k_cred_finish_any, k_chalpern_wr;
    hipFunction_t k_sc_minmax, k_sc_powsum, k_sc_factor, k_sc_apply, k_sc_mul, k_red_check, k_red_finish7;
    hipFunction_t k_tr_rows, k_tr_hist, k_scan_tiles, k_scan_add, k_tr_scatter, k_tr_ptr, k_tr_gather, k_tr_iota,
        k_remap, k_col_count, k_scale_inv_sqrt;
    hipFunction_t k_spmv, k_spmv_t_primal, k_spmv_t_primal_halpern, k_spmv_dual, k_spmv_dual_halpern;
    hipFunction_t k_primal_step, k_dual_step, k_halpern, k_sub, k_scale, k_fill, k_unscale_primal, k_unscale_dual;
    hipFunction_t k_probe_copy, k_probe_triad, k_probe_gather, k_probe_copy32, k_probe_triad32, k_replay64,
        k_replay32, k_probe_overlap;
    hipFunction_t k_spmv32, k_spmv_t_primal_mixed, k_spmv_dual_mixed, k_to_f32, k_delta_from, k_delta_to;
    // The A'ᵀ-side plain-iteration kernels with a hot table of y in LDS, and
    // the workgroups of each that one CU holds at once (their persistent grid
    // is that times the CUs).
    hipFunction_t k_spmv_t_primal_halpern_hot, k_spmv_t_primal_mixed_hot;
    int hot_occ64 = 0, hot_occ32 = 0;
    hipFunction_t k_panel_count, k_panel_gather, k_spmv_panel, k_fin_panel, k_spmv_panel32, k_fin_panel32;
    hipFunction_t k_rows_gather, k_gather;
    hipFunction_t k_dot, k_dist, k_primal_stats, k_dual_stats, k_step_dual, k_primal_ray, k_dual_ray;
};

namespace sp {

// Checks a HIP status and throws sp::Error(SP_ERR_HIP) with the call site.
void hip_check(hipError_t e, const char* what);

// The message of the calling thread's last failed call on backend `b` (""
// after a call that succeeded). Kept per thread, so a call that fails on one
// thread (a serving submit) never touches the text another thread reads.
void set_error(const sp_backend* b, const std::string& msg);
uint64_t next_backend_serial();
const char* last_error(const sp_backend* b);
void forget_errors(const sp_backend* b);

// The device address of `device`'s copy of the static scratch array of the
// no-allocation entry point (taken once per device, with `device` current),
// and its length in doubles (noalloc.hip).
double* scratch_on_device(int device);
int64_t scratch_doubles();

// The Backend methods. Each throws sp::Error on failure.
struct Hip {
    sp_backend* b;

    static constexpr int kMaxReduceBlocks = 1024;
    static constexpr int kMaxSlots = 8;
    static constexpr int kSlots = 64;

    std::string name() const { return b->name; }
    sp_problem* load(const ScaledProblem& s, bool autotune = false);
    sp_problem* load_view(const sp_scaled_view& v, const SpmvTuning* ta = nullptr, const SpmvTuning* tat = nullptr);
    // ScaledProblem::new on the device (scaling.rs:53-103): uploads A and its
    // transpose unscaled, runs the scaling passes there, and fills `s` with
    // the factors and scaled vectors (its matrices stay empty).
    // With `order`, the columns are relabeled by popularity on the device
    // after the upload (counts, then the order, then the column indices), and
    // `order` receives the relabeling.
    sp_problem* load_scaled(const LpView& p, const ScalingParams& sp, bool autotune, ScaledProblem& s,
                            ColumnOrder* order = nullptr);
    // The transpose of a device CSR matrix on the device (problem.rs:88-109):
    // T gets row pointers, column indices, values and its SpMV plan (the
    // autotuned rule's or the default); its row pointers are also returned
    // in `host_rp`.
    void transpose(const DevCsr& A, DevCsr& T, bool autotune, std::vector<int64_t>& host_rp);
    // Column panels for A' of a loaded problem: from now on every A' product
    // gathers the long rows' nonzeros panel by panel, 2^shift columns per
    // panel (shift in [8, 30]); shift 0 drops them. A matrix without long rows
    // keeps its plain plan.
    void set_panels(sp_problem* p, int shift);
    // Relabels the rows of a loaded problem (before its FP32 copy or panels
    // exist): row i of A' becomes row order[i], the row-indexed vectors
    // follow, and A'ᵀ's column indices are renamed in place.
    void permute_rows(sp_problem* p, const std::vector<int64_t>& order);
    // Decides whether A'ᵀ's plain-iteration kernels keep a hot table of y:
    // when most of its nonzeros (A' rows below H, so A'.rp[H] of them) gather
    // one of y's first H entries, as after long_rows_first on a matrix of
    // skewed row lengths (mode -1: and the matrix is large; 1: any size; 0:
    // never). Call after the last change of row order.
    void plan_hot(sp_problem* p, int mode);
    // Uploads a CSR matrix (int32 indices) and plans it, as load does.
    void upload(DevCsr& d, const int64_t* row_ptr, const int64_t* col_idx, const double* values, int64_t nrows,
                int64_t ncols, int64_t nnz, const SpmvTuning& tuning);
    // check() of crates/saddlepoint-check on the device, from the original LP.
    sp_kkt_report check(const LpView& p, const double* x, const double* y);
    // Average ms of one SpMV (A' or, with transpose, A'ᵀ) over reps launches.
    double spmv_time(const sp_problem* p, bool transpose, int reps);
    sp_vec* upload(const double* data, int64_t len);
    sp_vec* zeros(int64_t len);
    void download(const sp_vec* v, double* out, int64_t len);
    void copy(const sp_vec* src, sp_vec* dst);

    void spmv(const sp_problem* p, const sp_vec* x, sp_vec* out);
    void spmv_t(const sp_problem* p, const sp_vec* y, sp_vec* out);
    void primal_step(const sp_problem* p, const sp_vec* x, const sp_vec* aty, double tau, sp_vec* x_t,
                     sp_vec* x_bar);
    void dual_step(const sp_problem* p, const sp_vec* y, const sp_vec* ax_bar, double sigma, sp_vec* y_t);
    void halpern_step(sp_vec* z, const sp_vec* z_t, const sp_vec* anchor, double w, double rho);
    double distance_sq(const sp_vec* a, const sp_vec* bv);
    void sub(const sp_vec* a, const sp_vec* bv, sp_vec* out);
    // ScaledProblem::unscale_primal / unscale_dual (scaling.rs:114-121) on the
    // device: out = v · f / s entry by entry. With `to` (device int32, v's
    // length) entry q goes to out[to[q]], and the dual one takes f[to[q]].
    void unscale(const sp_vec* v, const double* f, int64_t flen, double s, const int* to, bool dual, sp_vec* out);
    // The same into host memory, for a vector in a relabeled order: the
    // factors and the relabeling (`order`, may be null or empty) are uploaded,
    // and the result comes back in the LP's order.
    void unscale_to_host(const sp_vec* v, const std::vector<double>& f, double s, const std::vector<int64_t>* order,
                         bool dual, double* out);
    double norm_sq(const sp_vec* v);
    double dot(const sp_vec* a, const sp_vec* bv);
    void scale(sp_vec* v, double alpha);
    sp_primal_stats_t primal_stats(const sp_problem* p, const sp_vec* x, const sp_vec* ax);
    sp_dual_stats_t dual_stats(const sp_problem* p, const sp_vec* y, const sp_vec* aty);
    // dual_stats with its partials and totals in caller-provided device memory
    // (`cap` doubles) instead of the backend's buffers; launches directly
    // (never through the profiler's events) and allocates nothing.
    sp_dual_stats_t dual_stats_in(const sp_problem* p, const sp_vec* y, const sp_vec* aty, double* ws, int64_t cap);
    double step_dual_residual_sq(const sp_problem* p, const sp_vec* x, const sp_vec* x_t, const sp_vec* aty,
                                 const sp_vec* aty_t, double tau);
    sp_ray_stats_t primal_ray_stats(const sp_problem* p, const sp_vec* dx, const sp_vec* adx);
    sp_ray_stats_t dual_ray_stats(const sp_problem* p, const sp_vec* dy, const sp_vec* atdy);

    // Fused forms used by the driver when params.fused is set. Each is
    // bitwise equal to the sequence of Backend calls it replaces.
    // spmv_t + primal_step
    void spmv_t_primal(const sp_problem* p, const sp_vec* y, const sp_vec* x, double tau, sp_vec* aty,
                       sp_vec* x_t, sp_vec* x_bar);
    // spmv + dual_step
    void spmv_dual(const sp_problem* p, const sp_vec* x_bar, const sp_vec* y, double sigma, sp_vec* ax,
                   sp_vec* y_t);
    // One whole iteration without a check: spmv_t + primal_step + halpern(x),
    // then spmv + dual_step + halpern(y). x and y are updated in place;
    // x_bar is overwritten; x_t, y_t, aty and ax are not written.
    void plain_iteration(const sp_problem* p, sp_vec* x, sp_vec* y, const sp_vec* x_anchor, const sp_vec* y_anchor,
                         sp_vec* x_bar, double tau, double sigma, double w, double rho);

    // Mixed precision (sp_params.precision = 1): iterates as an FP64 base (the
    // iterate at the last check) + an FP32 displacement between checks.
    void ensure_fp32(sp_problem* p);
    sp_vec32* zeros32(int64_t len);
    void delta_from(const sp_vec* z, const sp_vec* base, sp_vec32* dz);  // dz = z - base
    void delta_to(const sp_vec32* dz, const sp_vec* base, sp_vec* z);    // z = base + dz (base may be z)
    // plain_iteration in displacement form: x = x_base + dx, y = y_base + dy;
    // atya = A'ᵀy_base, axa = A'x_base; ex = x_anchor - x_base, ey = y_anchor - y_base.
    void plain_iteration_mixed(const sp_problem* p, sp_vec32* dx, sp_vec32* dy, const sp_vec* x_base,
                               const sp_vec* y_base, const sp_vec* atya, const sp_vec* axa, const sp_vec32* ex,
                               const sp_vec32* ey, sp_vec32* dx_bar, double tau, double sigma, double w, double rho);

    // Average milliseconds of one probe kernel over `reps` launches: kind 0
    // copies n doubles, 1 is a triad over n doubles, 2 gathers n doubles at
    // pseudo-random positions of a table of table_len doubles, 3 and 4 are
    // the copy and the triad over n floats; 5-8 are sp_probe_overlap modes
    // 0-3 (n gathers from such a table beside a stream of n 12-byte
    // elements: gathers only, stream only, alternate workgroups, both in
    // every workgroup), n rounded up to whole 2048-element blocks.
    double probe(int kind, int64_t n, int64_t table_len, int reps, uint64_t seed);
    // Average milliseconds of one launch of the gather-only replay of a
    // loaded matrix's column indices (A' or, with transpose, A'ᵀ): every
    // nonzero's entry of a table as long as the matrix has columns is loaded,
    // in the SpMV's order and tiling, and nothing else is read or written.
    // fp32: the table holds floats (the mixed-precision iterates).
    double gather_replay_time(const sp_problem* p, bool transpose, bool fp32, int reps);

    // The operands of the plain iteration's two fused kernels.
    struct PlainArgs {
        const sp_vec *x, *y, *x_anchor, *y_anchor, *x_bar;
        const sp_vec32 *dx, *dy, *dx_bar, *ex, *ey;  // mixed precision
        const sp_vec *atya, *axa;
        double tau, sigma, w, rho;
    };
    // Average milliseconds of one launch of a plain iteration's main fused
    // kernel (side 0: the A'ᵀ side, 1: the A' side; mixed: the FP32 forms),
    // from two events around `reps` back-to-back launches. The kernel writes
    // copies of the vectors it updates, so the operands are left unchanged.
    double plain_kernel_time(const sp_problem* p, int side, bool mixed, const PlainArgs& a, int reps);

    // Reductions into device result slots, read back together by fetch().
    double* slot(int i);
    const double* fetch(int count);
    // Reads the first `count` values of a device vector into the pinned
    // staging area (valid until the next fetch).
    const double* fetch_vec(const sp_vec* v, int count);
    // Enqueues the copy of the first `count` values of `v` into pinned host
    // memory `dst` (from host_alloc); valid after the next sync().
    void fetch_vec_async(const sp_vec* v, int count, double* dst);
    // Pinned host memory for fetch_vec_async.
    double* host_alloc(int64_t count);
    void dot_into(const sp_vec* a, const sp_vec* bv, double* dst);
    void distance_sq_into(const sp_vec* a, const sp_vec* bv, double* dst);
    void primal_stats_into(const sp_problem* p, const sp_vec* x, const sp_vec* ax, double* dst);      // 2 sums
    void dual_stats_into(const sp_problem* p, const sp_vec* y, const sp_vec* aty, double* dst);       // 3 sums
    void step_dual_residual_sq_into(const sp_problem* p, const sp_vec* x, const sp_vec* x_t, const sp_vec* aty,
                                    const sp_vec* aty_t, double tau, double* dst);
    void primal_ray_stats_into(const sp_problem* p, const sp_vec* dx, const sp_vec* adx, double* dst);  // 4 sums
    void dual_ray_stats_into(const sp_problem* p, const sp_vec* dy, const sp_vec* atdy, double* dst);   // 5 sums
    static sp_primal_stats_t primal_from(const double* raw);
    static sp_dual_stats_t dual_from(const double* raw);
    static sp_ray_stats_t primal_ray_from(const double* raw);
    static sp_ray_stats_t dual_ray_from(const double* raw);

    // v *= 1 / sqrt(*ns) with *ns on the device (a result slot).
    void scale_inv_sqrt(sp_vec* v, const double* ns);

    // An OpList on the device, and its three groups as launches: the
    // differences; the reductions (step_dual's entries with the given 1/τ,
    // rewritten in the table when it changes); the Halpern updates with the
    // given w and ρ.
    DevOpList* upload_ops(const OpList& l);
    void run_subs(const DevOpList* d);
    void run_reductions(DevOpList* d, double inv_tau);
    void run_halperns(const DevOpList* d, double w, double rho);

    void sync();

    // Workgroups of a reduction over n elements: a fixed function of n, so a
    // given reduction always sums in the same order.
    int reduce_grid(int64_t n) const;

   private:
    const double* read_result(int k, int grid);
    double probe_overlap(int mode, int64_t n, int64_t table_len, int reps, uint64_t seed);
};

void free_problem(sp_problem* p);
void free_vec(sp_vec* v);
void free_vec32(sp_vec32* v);
void free_csr_public(DevCsr& d);  // releases a DevCsr's device arrays
void free_host(double* h);
void free_ops(DevOpList* d);
// Folds finished event pairs into the per-kernel totals (synchronizes).
void profile_collect(sp_backend* b);

}  // namespace sp

// Opens device `ordinal` and loads the code object; throws sp::Error.
sp_backend* sp_backend_open(int ordinal);
void sp_backend_close(sp_backend* b);
