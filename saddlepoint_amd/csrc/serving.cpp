// Continuous-batching serving loop on one GPU.
//
// LPs are submitted at any time. Host worker threads validate them and scale
// the small ones (ScaledProblem::new on the host; large ones are scaled on the
// device at admission); the loop admits prepared LPs while the HBM budget and
// the slot count allow, keeps every admitted LP's iterates, anchors, ω, step
// sizes and restart state on the device, and retires LPs as they terminate.
// Each chunk runs the admitted LPs that the chunk schedule selects (the
// cheapest up to a work target, and any LP left out of the last chunks).
//
// The loop runs in chunks of 64 iterations aligned to every LP's own check
// cadence (an LP is admitted at a chunk boundary with its iteration count at
// 0, so its checks fall on the chunk's last step). In a chunk:
//   step 1      LPs that start a restart epoch (k = 0) take the iteration
//               that also measures the anchor's fixed-point residual, as one
//               batched phase; the others take a batched plain iteration;
//   steps 2-63  one batched plain iteration for all LPs per step: one
//               launch per side, over the concatenated work items of every
//               LP's own SpMV plan (optionally replayed from a hipGraph);
//   step 64     every LP's check (KKT, infeasibility, restart, PID) in FP64:
//               the device work as one batched phase (batch_check.hpp), one
//               readback of every LP's result slots, then each LP's
//               decisions by the code of a single solve.
// Because each LP keeps its own plan and its own per-LP reductions, every
// LP's iterates are bitwise the same alone, in any mix, and in a single
// solve.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "batch_check.hpp"
#include "driver.hpp"
#include "saddlepoint_hip.h"
#include "solver.hpp"

namespace sp {

namespace {

// Host mirrors of BInst / BDesc / BLong in kernels.hip (same layout).
struct BInst {
    const int *a_rp, *a_ci;
    const double* a_va;
    const float* a_va32;
    double* a_part;
    const int *t_rp, *t_ci;
    const double* t_va;
    const float* t_va32;
    double* t_part;
    const double *c, *lc, *uc, *lv, *uv;
    double *x, *y, *x_bar;
    const double *x_anchor, *y_anchor;
    float *dx, *dy, *dx_bar;
    const double *atya, *axa;
    const float *ex, *ey;
    double tau, sigma, rho;
    unsigned long long k0;
};
static_assert(sizeof(BInst) == 248, "BInst layout is shared with kernels.hip");
using Clock = std::chrono::steady_clock;

constexpr uint64_t kChunk = 64;
// LPs with fewer nonzeros are scaled by the prep threads on the host.
constexpr int64_t kHostScalingBelow = 2'000'000;

template <class T>
T* dalloc(size_t n) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
    return static_cast<T*>(p);
}

// A growable device buffer.
template <class T>
struct DevBuf {
    T* d = nullptr;
    size_t cap = 0;
    void ensure(size_t n) {
        if (n <= cap) return;
        if (d) (void)hipFree(d);
        cap = std::max(n, cap * 2);
        d = dalloc<T>(cap);
    }
    ~DevBuf() {
        if (d) (void)hipFree(d);
    }
};

}  // namespace

enum JobState { kQueued = 0, kPrepared = 1, kActive = 2, kDone = 3, kFailed = 4 };

struct Job {
    uint64_t id = 0;
    int64_t m = 0, n = 0;
    std::vector<int64_t> rp, ci;
    std::vector<double> va, c, lc, uc, lv, uv;
    double c0 = 0.0;
    bool maximize = false;
    sp_params params{};
    double t_submit = 0, t_admit = -1, t_done = -1;
    int state = kQueued;
    std::unique_ptr<ScaledProblem> scaled;
    ColumnOrder cols;  // the relabeling already applied to `scaled` (host scaling)
    std::unique_ptr<Solver<DriverBackend>> solver;
    int64_t footprint = 0;
    sp_solution sol{};
    std::vector<double> x, y;
    int err = SP_OK;
    std::string msg;
    bool failed_now = false;  // an error in this chunk; retired at its end
    int64_t cost = 0;         // per-iteration work, for the chunk schedule: nnz + m + n
    int skipped = 0;          // chunks in a row this LP was left out of

    LpView view() const {
        LpView v;
        v.m = m;
        v.n = n;
        v.nnz = (int64_t)va.size();
        v.row_ptr = rp.data();
        v.col_idx = ci.data();
        v.values = va.data();
        v.c = c.data();
        v.lc = lc.data();
        v.uc = uc.data();
        v.lv = lv.data();
        v.uv = uv.data();
        v.c0 = c0;
        v.maximize = maximize;
        return v;
    }
};

// Device bytes one admitted LP holds (the admission test): both matrices with
// int32 indices and FP64 values (+ FP32 values in mixed precision), the
// solver's 13 vectors and the 6 of its check iterations, the problem's 7
// vectors and the mixed-precision displacements, anchor offsets and base
// products.
static int64_t footprint(int64_t m, int64_t n, int64_t nnz, bool mixed) {
    int64_t bytes = 2 * (12 * nnz + 4 * (m + 1)) + 8 * (10 * n + 9 * m) + 8 * (4 * n + 3 * m);
    bytes += 2 * 32 * (nnz / 1024 + 64);  // plans
    if (mixed) bytes += 2 * 4 * nnz + 4 * (3 * n + 2 * m) + 8 * (n + m);
    return bytes;
}

class Server {
   public:
    Server(sp_backend* b, const sp_server_params& sp) : b_(b), db_{{b}}, sp_(sp), t0_(Clock::now()) {
        const int workers = std::max(1, sp_.prep_threads);
        for (int i = 0; i < workers; ++i) workers_.emplace_back([this] { prep_loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
        (void)hipStreamSynchronize(b_->stream);
        for (auto& kv : jobs_) kv.second->solver.reset();
        if (h_inst_) (void)hipHostFree(h_inst_);
        if (cpart_) (void)hipFree(cpart_);
        if (gath_d_) (void)hipFree(gath_d_);
        if (gath_h_) (void)hipHostFree(gath_h_);
    }

    double clock() const { return std::chrono::duration<double>(Clock::now() - t0_).count(); }

    uint64_t submit(const sp_lp_view& lp, const sp_params& params) {
        auto job = std::make_unique<Job>();
        job->m = lp.a.nrows;
        job->n = lp.a.ncols;
        const int64_t nnz = lp.a.nnz;
        if (job->m < 0 || job->n < 0 || nnz < 0 || !lp.a.row_ptr) throw Error(SP_ERR_INVALID, "submit: bad shape");
        job->rp.assign(lp.a.row_ptr, lp.a.row_ptr + job->m + 1);
        job->ci.assign(lp.a.col_idx, lp.a.col_idx + nnz);
        job->va.assign(lp.a.values, lp.a.values + nnz);
        job->c.assign(lp.c, lp.c + job->n);
        job->lv.assign(lp.lv, lp.lv + job->n);
        job->uv.assign(lp.uv, lp.uv + job->n);
        job->lc.assign(lp.lc, lp.lc + job->m);
        job->uc.assign(lp.uc, lp.uc + job->m);
        job->c0 = lp.c0;
        job->maximize = lp.maximize != 0;
        job->params = params;
        // The loop's chunks follow the reference's check cadence of 64; the
        // iteration limit is rounded up to a whole number of chunks.
        if (params.check_every != kChunk) throw Error(SP_ERR_INVALID, "serving needs check_every = 64");
        job->params.fused = 1;
        // The batched kernels run each LP's plain plan, so its single-solve
        // counterpart is the one without column panels.
        job->params.column_panels = 0;
        // Limits past the largest multiple of 64 (u64::MAX, "no limit") become it.
        const uint64_t last = UINT64_MAX / kChunk * kChunk;
        job->params.max_iterations =
            params.max_iterations > last ? last : (params.max_iterations + kChunk - 1) / kChunk * kChunk;
        job->footprint = footprint(job->m, job->n, nnz, params.precision == 1);
        job->cost = nnz + job->m + job->n;
        job->t_submit = clock();
        std::lock_guard<std::mutex> g(mu_);
        job->id = ++next_id_;
        const uint64_t id = job->id;
        to_prep_.push_back(job.get());
        jobs_[id] = std::move(job);
        cv_.notify_one();
        return id;
    }

    // Runs one 64-iteration chunk for every admitted LP (admitting and
    // retiring at the chunk's edges).
    void run_chunk(int* active, int* waiting, int* done_total) {
        admit();
        if (!active_.empty()) chunk();
        retire();
        if (active) *active = (int)active_.size();
        std::lock_guard<std::mutex> g(mu_);
        if (waiting) *waiting = (int)(to_prep_.size() + ready_.size());
        if (done_total) *done_total = (int)finished_;
    }

    int state(uint64_t id) {
        std::lock_guard<std::mutex> g(mu_);
        auto it = jobs_.find(id);
        if (it == jobs_.end()) throw Error(SP_ERR_INVALID, "no such job");
        return it->second->state;
    }

    void result(uint64_t id, sp_solution* out, double* x, double* y, sp_job_times* t) {
        std::lock_guard<std::mutex> g(mu_);
        auto it = jobs_.find(id);
        if (it == jobs_.end()) throw Error(SP_ERR_INVALID, "no such job");
        Job& j = *it->second;
        if (j.state == kFailed) throw Error(j.err, j.msg);
        if (j.state != kDone) throw Error(SP_ERR_INVALID, "job not finished");
        if (out) *out = j.sol;
        if (x) std::copy(j.x.begin(), j.x.end(), x);
        if (y) std::copy(j.y.begin(), j.y.end(), y);
        if (t) *t = sp_job_times{j.t_submit, j.t_admit, j.t_done};
    }

    void forget(uint64_t id) {
        std::lock_guard<std::mutex> g(mu_);
        auto it = jobs_.find(id);
        if (it != jobs_.end() && (it->second->state == kDone || it->second->state == kFailed)) jobs_.erase(it);
    }

   private:
    void prep_loop() {
        for (;;) {
            Job* j;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !to_prep_.empty(); });
                if (stop_) return;
                j = to_prep_.front();
                to_prep_.pop_front();
            }
            try {
                const LpView v = j->view();
                validate(v);
                // Small LPs are scaled (and relabeled) here, off the GPU's
                // stream; large ones by the device at admission. Both give
                // bitwise the same scaled problem.
                if (j->params.gpu_scaling && v.nnz < kHostScalingBelow) j->params.gpu_scaling = 0;
                if (j->params.gpu_scaling) {
                    j->scaled = std::make_unique<ScaledProblem>();
                } else {
                    j->scaled = std::make_unique<ScaledProblem>(ScaledProblem::build(v, scaling_params_of(j->params)));
                    if (j->params.reorder_columns) {
                        j->cols = popularity_order(v);
                        relabel(*j->scaled, j->cols);
                    }
                }
                std::lock_guard<std::mutex> g(mu_);
                j->state = kPrepared;
                ready_.push_back(j);
            } catch (const Error& e) {
                fail(j, e.code, e.what());
            } catch (const std::exception& e) {
                fail(j, SP_ERR_INVALID, e.what());
            }
        }
    }

    void fail(Job* j, int code, const std::string& msg) {
        std::lock_guard<std::mutex> g(mu_);
        j->err = code;
        j->msg = msg;
        j->state = kFailed;
        j->t_done = clock();
        ++finished_;
    }

    void admit() {
        // Take LPs while the budget and the slots allow; their power
        // iterations then run together.
        std::vector<Job*> fresh;
        for (;;) {
            Job* j = nullptr;
            {
                std::lock_guard<std::mutex> g(mu_);
                if (ready_.empty()) break;
                Job* f = ready_.front();
                const bool empty = active_.empty() && fresh.empty();
                const bool fits = used_ + f->footprint <= sp_.budget_bytes || empty;
                if ((int)(active_.size() + fresh.size()) >= sp_.max_active || !fits) break;
                ready_.pop_front();
                j = f;
            }
            try {
                j->solver = std::make_unique<Solver<DriverBackend>>(db_, j->view(), j->params, std::move(*j->scaled),
                                                                    std::move(j->cols));
                j->scaled.reset();
                j->solver->power_begin();
                used_ += j->footprint;
                fresh.push_back(j);
            } catch (const Error& e) {
                j->solver.reset();
                fail(j, e.code, e.what());
            }
        }
        if (fresh.empty()) return;
        power_iterations(fresh);
        for (Job* j : fresh) {
            if (!j->failed_now) {
                try {
                    j->solver->power_end();
                } catch (const Error& e) {
                    mark_failed(j, e);
                }
            }
            if (j->failed_now) {
                j->solver.reset();
                used_ -= j->footprint;
                fail(j, j->err, j->msg);
                continue;
            }
            j->t_admit = clock();
            {
                std::lock_guard<std::mutex> g(mu_);
                j->state = kActive;
            }
            active_.push_back(j);
        }
    }

    // The power iterations of newly admitted LPs, batched: blocks of rounds
    // with one readback of every LP's norms per block (Solver::start's
    // arithmetic, so each LP's step size is the one of its single solve).
    void power_iterations(const std::vector<Job*>& fresh) {
        constexpr int K = Solver<DriverBackend>::power_block();
        constexpr int base = Solver<DriverBackend>::slot_power();
        for (;;) {
            std::vector<Job*> live;
            std::vector<PowerWork> ws;
            int rounds = 0;
            for (Job* j : fresh) {
                if (j->failed_now || j->solver->power_done()) continue;
                const int r = j->solver->power_rounds_next();
                if (live.empty()) rounds = r;
                if (r != rounds) continue;  // a different cap: its own block later
                live.push_back(j);
                ws.push_back(j->solver->power_work());
            }
            if (live.empty()) return;
            ensure_gather(live.size() * K);
            Phase P(b_);
            enqueue_power_rounds(P, db_, ws, rounds);
            std::vector<COp> copy;
            for (size_t i = 0; i < ws.size(); ++i) {
                COp o;
                o.a0 = ws[i].slots + base, o.o0 = gath_d_ + i * K, o.n = rounds;
                copy.push_back(o);
            }
            P.elementwise(b_->k_ccopy, copy);
            P.run(arena_, cpart_, cpart_cap_);
            hip_check(hipMemcpyAsync(gath_h_, gath_d_, ws.size() * K * sizeof(double), hipMemcpyDeviceToHost,
                                     b_->stream),
                      "hipMemcpyAsync D2H");
            hip_check(hipStreamSynchronize(b_->stream), "hipStreamSynchronize");
            for (size_t i = 0; i < live.size(); ++i) {
                try {
                    live[i]->solver->power_take(gath_h_ + i * K, rounds);
                } catch (const Error& e) {
                    mark_failed(live[i], e);
                }
            }
        }
    }

    // Runs one LP's own iteration (the check path of a single solve).
    void own_step(Job* j) {
        try {
            j->solver->advance(1);
        } catch (const Error& e) {
            j->err = e.code;
            j->msg = e.what();
            j->failed_now = true;
        }
    }

    // The LPs this chunk runs: the cheapest first (cost, then id) until their
    // costs reach the chunk's work target, and every LP that was left out of
    // the last max_skipped chunks; in id order. The rest keep their state.
    static constexpr int64_t kChunkWork = 32'000'000;
    static constexpr int kMaxSkipped = 8;
    std::vector<Job*> chunk_members() {
        std::vector<Job*> live;
        for (Job* j : active_)
            if (!j->failed_now && !j->solver->terminated()) live.push_back(j);
        if (sp_.chunk_work < 0) return live;
        const int64_t target = sp_.chunk_work > 0 ? sp_.chunk_work : kChunkWork;
        const int max_skipped = sp_.max_skipped > 0 ? sp_.max_skipped : kMaxSkipped;
        std::vector<Job*> order = live;
        std::stable_sort(order.begin(), order.end(), [](const Job* a, const Job* b) {
            return a->cost != b->cost ? a->cost < b->cost : a->id < b->id;
        });
        int64_t work = 0;
        std::vector<Job*> run;
        for (Job* j : order) {
            if (work < target || j->skipped >= max_skipped) {
                run.push_back(j);
                work += j->cost;
                j->skipped = 0;
            } else {
                ++j->skipped;
            }
        }
        std::sort(run.begin(), run.end(), [](const Job* a, const Job* b) { return a->id < b->id; });
        return run;
    }

    void chunk() {
        // Everything of the previous chunk has left the stream (its last
        // work, the restart copies, is waited for here), so the pinned
        // staging areas can be reused.
        hip_check(hipStreamSynchronize(b_->stream), "hipStreamSynchronize");
        arena_.reset();
        staged_ = 0;
        ensure_staging(2 * active_.size());
        const std::vector<Job*> members = chunk_members();
        // Step 1.
        std::vector<Job*> first, batch1;
        for (Job* j : members) {
            if (j->failed_now || j->solver->terminated()) continue;
            if (j->solver->plain_run() == 0) {
                if (!j->solver->check_due()) first.push_back(j);
            } else {
                batch1.push_back(j);
            }
        }
        run_steps(first);
        if (!batch1.empty()) run_batched(batch1, 1, false, first64_, first32_);
        // Steps 2-63: everyone that is still running and in step.
        std::vector<Job*> all;
        uint64_t steps = UINT64_MAX;
        for (Job* j : members) {
            if (j->failed_now || j->solver->terminated()) continue;
            all.push_back(j);
            steps = std::min(steps, j->solver->plain_run());
        }
        if (!all.empty() && steps > 0 && steps != UINT64_MAX)
            run_batched(all, (int)steps, sp_.use_graphs != 0, batch64_, batch32_);
        // Step 64: every LP's check.
        std::vector<Job*> due;
        for (Job* j : members) {
            while (!j->failed_now && !j->solver->terminated() && j->solver->plain_run() == 0 &&
                   !j->solver->check_due())
                own_step(j);  // a restart epoch's first iteration (not reached with aligned chunks)
            if (!j->failed_now && j->solver->check_due()) due.push_back(j);
        }
        run_steps(due);
    }

    // The next non-plain iteration of every LP in `group` (all checks, or
    // all first iterations of a restart epoch) as one batched phase; for
    // checks, one readback and each LP's decisions, then the restart copies.
    void run_steps(const std::vector<Job*>& group) {
        if (group.empty()) return;
        std::vector<StepWork> ws;
        std::vector<Job*> who;
        for (Job* j : group) {
            try {
                ws.push_back(j->solver->begin_step());
                who.push_back(j);
            } catch (const Error& e) {
                mark_failed(j, e);
            }
        }
        if (ws.empty()) return;
        const bool check = ws[0].check;
        for (const StepWork& w : ws)
            if (w.check != check) throw Error(SP_ERR_INVALID, "serving: mixed step kinds in one phase");
        constexpr int K = Solver<DriverBackend>::slot_count();
        Phase P(b_);
        enqueue_step_work(P, db_, ws);
        if (check) {
            ensure_gather(ws.size() * K);
            gather_slots(P, db_, ws, gath_d_, K);
        }
        P.run(arena_, cpart_, cpart_cap_);
        if (!check) return;
        hip_check(hipMemcpyAsync(gath_h_, gath_d_, ws.size() * K * sizeof(double), hipMemcpyDeviceToHost,
                                 b_->stream),
                  "hipMemcpyAsync D2H");
        hip_check(hipStreamSynchronize(b_->stream), "hipStreamSynchronize");
        std::vector<std::pair<sp_vec*, sp_vec*>> copies;
        for (size_t i = 0; i < who.size(); ++i) {
            try {
                who[i]->solver->end_step(gath_h_ + i * K, &copies);
            } catch (const Error& e) {
                mark_failed(who[i], e);
            }
        }
        Phase C(b_);
        enqueue_copies(C, db_, copies);
        C.run(arena_, cpart_, cpart_cap_);
    }

    void mark_failed(Job* j, const Error& e) {
        j->err = e.code;
        j->msg = e.what();
        j->failed_now = true;
    }

    void ensure_gather(size_t n) {
        if (gath_cap_ >= n) return;
        hip_check(hipStreamSynchronize(b_->stream), "hipStreamSynchronize");
        if (gath_h_) (void)hipHostFree(gath_h_);
        if (gath_d_) (void)hipFree(gath_d_);
        gath_h_ = nullptr;
        gath_d_ = nullptr;
        gath_cap_ = std::max(n, 2 * gath_cap_);
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&gath_h_), gath_cap_ * sizeof(double), hipHostMallocDefault),
                  "hipHostMalloc");
        hip_check(hipMalloc(reinterpret_cast<void**>(&gath_d_), gath_cap_ * sizeof(double)), "hipMalloc");
    }

    struct Batch;

    // `steps` batched plain iterations of `group`, through the FP64 and the
    // mixed-precision batch tables.
    void run_batched(const std::vector<Job*>& group, int steps, bool graph, Batch& b64, Batch& b32) {
        // Mixed-precision LPs entering plain iterations: the base's products,
        // zero displacements and the anchors' offsets, for all of them at once.
        std::vector<PlainEntry> entries;
        for (Job* j : group) {
            PlainEntry e;
            if (j->solver->begin_plain(e)) entries.push_back(e);
        }
        if (!entries.empty()) {
            Phase P(b_);
            enqueue_plain_entries(P, db_, entries);
            P.run(arena_, cpart_, cpart_cap_);
        }
        std::vector<Job*> g64, g32;
        for (Job* j : group) (j->solver->mixed_now() ? g32 : g64).push_back(j);
        for (int mixed = 0; mixed < 2; ++mixed) {
            const std::vector<Job*>& g = mixed ? g32 : g64;
            if (g.empty()) continue;
            Batch& bt = mixed ? b32 : b64;
            const bool same = build(bt, g, mixed != 0);
            // A graph pays off only when the membership carries over: a new
            // one is captured on the second chunk of a membership.
            const bool use_graph = graph && steps > 1 && same;
            launch_steps(bt, steps, mixed != 0, use_graph);
            for (Job* j : g) j->solver->plain_done((uint64_t)steps);
        }
    }

    struct Batch {
        std::vector<uint64_t> members;  // job ids (a freed job's address can come back)
        // The concatenated work lists (host copies live until the next
        // rebuild, so that their uploads need no synchronization).
        std::vector<BDesc> h_dt, h_da;
        std::vector<BLong> h_lt, h_la;
        DevBuf<BDesc> desc_t, desc_a;
        DevBuf<BLong> long_t, long_a;
        DevBuf<BInst> inst;
        int nt = 0, na = 0, nlt = 0, nla = 0;
        // The captured steps and their executable; the template graph is
        // kept as long as the executable (tools that trace graph launches
        // may look its nodes up).
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int graph_steps = 0;
        void drop_graph() {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            exec = nullptr;
            graph = nullptr;
        }
        ~Batch() { drop_graph(); }
    };

    void ensure_staging(size_t n) {
        if (host_inst_cap_ >= n) return;
        hip_check(hipStreamSynchronize(b_->stream), "hipStreamSynchronize");
        if (h_inst_) (void)hipHostFree(h_inst_);
        h_inst_ = nullptr;
        host_inst_cap_ = std::max<size_t>(n, 2 * host_inst_cap_);
        hip_check(
            hipHostMalloc(reinterpret_cast<void**>(&h_inst_), host_inst_cap_ * sizeof(BInst), hipHostMallocDefault),
            "hipHostMalloc");
    }

    // Returns whether the membership is that of the previous build.
    bool build(Batch& bt, const std::vector<Job*>& g, bool mixed) {
        std::vector<uint64_t> ids;
        for (Job* j : g) ids.push_back(j->id);
        const bool same = bt.members == ids;
        if (!same) {
            // The previous uploads from these host vectors were ordered on the
            // stream before any work of this chunk; a check synchronized since.
            bt.h_dt.clear();
            bt.h_da.clear();
            bt.h_lt.clear();
            bt.h_la.clear();
            for (size_t i = 0; i < g.size(); ++i) {
                const sp_problem* p = g[i]->solver->plain_state().problem;
                for (const Desc& d : p->at.host_desc) bt.h_dt.push_back(BDesc{d, (int)i, 0, 0, 0});
                for (const Desc& d : p->a.host_desc) bt.h_da.push_back(BDesc{d, (int)i, 0, 0, 0});
                for (const LongRow& l : p->at.host_longs) bt.h_lt.push_back(BLong{(int)i, l.row, l.pbase, l.nch});
                for (const LongRow& l : p->a.host_longs) bt.h_la.push_back(BLong{(int)i, l.row, l.pbase, l.nch});
            }
            auto up = [&](auto& buf, const auto& v) {
                buf.ensure(v.size());
                if (!v.empty())
                    hip_check(hipMemcpyAsync(buf.d, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice,
                                             b_->stream),
                              "hipMemcpyAsync H2D");
            };
            bt.drop_graph();
            up(bt.desc_t, bt.h_dt);
            up(bt.desc_a, bt.h_da);
            up(bt.long_t, bt.h_lt);
            up(bt.long_a, bt.h_la);
            bt.nt = (int)bt.h_dt.size();
            bt.na = (int)bt.h_da.size();
            bt.nlt = (int)bt.h_lt.size();
            bt.nla = (int)bt.h_la.size();
            bt.members = std::move(ids);
            bt.inst.ensure(g.size());
        }
        // The per-LP table: arrays and this step's parameters.
        if (staged_ + g.size() > host_inst_cap_) throw Error(SP_ERR_INVALID, "serving: staging area overflow");
        BInst* stage = h_inst_ + staged_;
        staged_ += g.size();
        for (size_t i = 0; i < g.size(); ++i) {
            const auto s = g[i]->solver->plain_state();
            const sp_problem* p = s.problem;
            BInst& I = stage[i];
            std::memset(&I, 0, sizeof I);
            I.a_rp = p->a.rp;
            I.a_ci = p->a.ci;
            I.a_va = p->a.va;
            I.a_va32 = p->a.va32;
            I.a_part = p->a.partials;
            I.t_rp = p->at.rp;
            I.t_ci = p->at.ci;
            I.t_va = p->at.va;
            I.t_va32 = p->at.va32;
            I.t_part = p->at.partials;
            I.c = p->c;
            I.lc = p->lc;
            I.uc = p->uc;
            I.lv = p->lv;
            I.uv = p->uv;
            I.x = s.x->d;
            I.y = s.y->d;
            I.x_bar = s.x_bar->d;
            I.x_anchor = s.x_anchor->d;
            I.y_anchor = s.y_anchor->d;
            if (mixed) {
                I.dx = s.dx->d;
                I.dy = s.dy->d;
                I.dx_bar = s.dx_bar->d;
                I.atya = s.atya->d;
                I.axa = s.axa->d;
                I.ex = s.ex->d;
                I.ey = s.ey->d;
            }
            I.tau = s.tau;
            I.sigma = s.sigma;
            I.rho = s.rho;
            I.k0 = s.k;
        }
        hip_check(hipMemcpyAsync(bt.inst.d, stage, g.size() * sizeof(BInst), hipMemcpyHostToDevice, b_->stream),
                  "hipMemcpyAsync H2D");
        return same;
    }

    template <class... Args>
    void kl(hipFunction_t f, unsigned grid, Args... args) {
        void* params[] = {static_cast<void*>(&args)...};
        hip_check(hipModuleLaunchKernel(f, grid, 1, 1, 256, 1, 1, 0, b_->stream, params, nullptr),
                  "hipModuleLaunchKernel");
    }

    void enqueue_steps(const Batch& bt, int steps, bool mixed) {
        hipFunction_t kt = mixed ? b_->k_bspmv_t_mixed : b_->k_bspmv_t64;
        hipFunction_t ka = mixed ? b_->k_bspmv_a_mixed : b_->k_bspmv_a64;
        hipFunction_t ft = mixed ? b_->k_bfin_t_mixed : b_->k_bfin_t64;
        hipFunction_t fa = mixed ? b_->k_bfin_a_mixed : b_->k_bfin_a64;
        for (int it = 0; it < steps; ++it) {
            if (bt.nt) kl(kt, (unsigned)bt.nt, (const BDesc*)bt.desc_t.d, (const BInst*)bt.inst.d, it);
            if (bt.nlt) kl(ft, (unsigned)((bt.nlt + 3) / 4), (const BLong*)bt.long_t.d, bt.nlt, (const BInst*)bt.inst.d, it);
            if (bt.na) kl(ka, (unsigned)bt.na, (const BDesc*)bt.desc_a.d, (const BInst*)bt.inst.d, it);
            if (bt.nla) kl(fa, (unsigned)((bt.nla + 3) / 4), (const BLong*)bt.long_a.d, bt.nla, (const BInst*)bt.inst.d, it);
        }
    }

    void launch_steps(Batch& bt, int steps, bool mixed, bool graph) {
        if (!graph) {
            enqueue_steps(bt, steps, mixed);
            return;
        }
        if (!bt.exec || bt.graph_steps != steps) {
            bt.drop_graph();
            hipGraph_t graph_ = nullptr;
            hip_check(hipStreamBeginCapture(b_->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                enqueue_steps(bt, steps, mixed);
            } catch (...) {
                (void)hipStreamEndCapture(b_->stream, &graph_);
                if (graph_) (void)hipGraphDestroy(graph_);
                throw;
            }
            hip_check(hipStreamEndCapture(b_->stream, &graph_), "hipStreamEndCapture");
            bt.graph = graph_;
            hip_check(hipGraphInstantiate(&bt.exec, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
            bt.graph_steps = steps;
        }
        hip_check(hipGraphLaunch(bt.exec, b_->stream), "hipGraphLaunch");
    }

    void retire() {
        std::vector<Job*> keep;
        for (Job* j : active_) {
            if (!j->failed_now && !j->solver->terminated()) {
                keep.push_back(j);
                continue;
            }
            if (j->failed_now) {
                j->solver.reset();
                fail(j, j->err, j->msg);
            } else {
                j->x.resize((size_t)j->n);
                j->y.resize((size_t)j->m);
                try {
                    j->sol = j->solver->finish(j->x.data(), j->y.data());
                    j->sol.setup_time_s = j->t_admit - j->t_submit;
                    j->solver.reset();
                    std::lock_guard<std::mutex> g(mu_);
                    j->t_done = clock();
                    j->sol.total_time_s = j->t_done - j->t_submit;
                    j->state = kDone;
                    ++finished_;
                } catch (const Error& e) {
                    j->solver.reset();
                    fail(j, e.code, e.what());
                }
            }
            used_ -= j->footprint;
        }
        active_.swap(keep);
    }

    sp_backend* b_;
    DriverBackend db_;
    sp_server_params sp_;
    Clock::time_point t0_;
    std::mutex mu_;
    std::condition_variable cv_;
    bool stop_ = false;
    std::deque<Job*> to_prep_, ready_;
    std::vector<std::thread> workers_;
    std::map<uint64_t, std::unique_ptr<Job>> jobs_;
    uint64_t next_id_ = 0;
    uint64_t finished_ = 0;
    std::vector<Job*> active_;
    int64_t used_ = 0;
    Batch first64_, first32_;  // step 1 of a chunk (membership changes with every restart)
    Batch batch64_, batch32_;  // steps 2-63 (membership changes with admissions and retirements)
    BInst* h_inst_ = nullptr;
    size_t host_inst_cap_ = 0;
    size_t staged_ = 0;
    // Batched phases: table arena, reduction partials, the result slots'
    // gather buffer and its pinned host copy.
    TableArena arena_{b_};
    double* cpart_ = nullptr;
    size_t cpart_cap_ = 0;
    double* gath_d_ = nullptr;
    double* gath_h_ = nullptr;
    size_t gath_cap_ = 0;
};

}  // namespace sp

struct sp_server {
    sp_backend* b;
    std::unique_ptr<sp::Server> s;
};

namespace {

template <class F>
int guarded_server(sp_server* sv, F&& f) {
    if (!sv) return SP_ERR_INVALID;
    try {
        f();
        sp::set_error(sv->b, "");
        return SP_OK;
    } catch (const sp::Error& e) {
        sp::set_error(sv->b, e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        sp::set_error(sv->b, "host out of memory");
        return SP_ERR_MEMORY;
    } catch (const std::exception& e) {
        sp::set_error(sv->b, e.what());
        return SP_ERR_INVALID;
    }
}

}  // namespace

extern "C" {

void sp_default_server_params(sp_server_params* p) {
    if (!p) return;
    p->budget_bytes = (int64_t)64 << 30;
    p->max_active = 256;
    p->prep_threads = 8;
    p->use_graphs = 1;
    p->chunk_work = 0;
    p->max_skipped = 0;
}

int sp_server_create(sp_backend* b, const sp_server_params* params, sp_server** out) {
    if (!b || !out) return SP_ERR_INVALID;
    *out = nullptr;
    sp_server_params p;
    sp_default_server_params(&p);
    if (params) p = *params;
    try {
        auto* sv = new sp_server{b, std::make_unique<sp::Server>(b, p)};
        *out = sv;
        return SP_OK;
    } catch (const std::exception& e) {
        sp::set_error(b, e.what());
        return SP_ERR_INVALID;
    }
}

void sp_server_free(sp_server* s) { delete s; }

int sp_server_submit(sp_server* s, const sp_lp_view* lp, const sp_params* params, uint64_t* id) {
    return guarded_server(s, [&] {
        if (!lp || !params || !id) throw sp::Error(SP_ERR_INVALID, "submit: null argument");
        *id = s->s->submit(*lp, *params);
    });
}

int sp_server_run_chunk(sp_server* s, int* active, int* waiting, int* finished) {
    return guarded_server(s, [&] { s->s->run_chunk(active, waiting, finished); });
}

int sp_server_state(sp_server* s, uint64_t id, int* state) {
    return guarded_server(s, [&] {
        if (!state) throw sp::Error(SP_ERR_INVALID, "state: null output");
        *state = s->s->state(id);
    });
}

int sp_server_result(sp_server* s, uint64_t id, sp_solution* out, double* x, double* y, sp_job_times* times) {
    return guarded_server(s, [&] { s->s->result(id, out, x, y, times); });
}

int sp_server_forget(sp_server* s, uint64_t id) {
    return guarded_server(s, [&] { s->s->forget(id); });
}

double sp_server_clock(const sp_server* s) { return s ? s->s->clock() : 0.0; }

}  // extern "C"
