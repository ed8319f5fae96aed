// The PDHG driver: restarted reflected-Halpern PDHG with the PID primal-weight
// controller, restated from crates/saddlepoint-core/src/solver.rs:154-480 so
// that it drives any backend with the reference's Backend method sequence.
//
// Templated on the backend. With params.fused, the iterations between checks
// run through the backend's fused kernels instead (B::plain_iteration, and
// B::spmv_t_primal / B::spmv_dual on check iterations), which compute the same
// expressions in the same order and so give bitwise the same iterates.
#pragma once

#include <chrono>
#include <cmath>
#include <algorithm>
#include <cstdio>
#include <limits>
#include <memory>
#include <utility>
#include <vector>

#include "host.hpp"
#include "saddlepoint_hip.h"

namespace sp {

// ScalingParams from the C ABI's parameters (scaling.rs:25-34).
inline ScalingParams scaling_params_of(const sp_params& p) {
    ScalingParams s;
    s.geometric_mean_passes = p.geometric_mean_passes;
    s.ruiz_iterations = p.ruiz_iterations;
    s.has_pock_chambolle = p.has_pock_chambolle != 0;
    s.pock_chambolle_alpha = p.pock_chambolle_alpha;
    s.bound_objective_rescaling = p.bound_objective_rescaling != 0;
    return s;
}

// PidState (solver.rs:445-480).
struct PidState {
    double integral = 0.0;
    bool has_prev = false;
    double prev_error = 0.0;
    double best_weight;
    double best_balance = std::numeric_limits<double>::infinity();

    explicit PidState(double omega) : best_weight(omega) {}

    double update(const sp_params& p, double omega, double dx, double dy, double rel_primal, double rel_dual) {
        auto sane = [](double v) { return v >= 1e-16 && v <= 1e12; };
        const double ratio = rel_dual / rel_primal;
        if (!sane(dx) || !sane(dy) || !(ratio >= 1e-8 && ratio <= 1e8)) {
            integral = 0.0;
            has_prev = false;
            return best_weight;
        }
        // Positive error: the scaled primal displacement is too large.
        const double e = std::log(omega * dx / dy);
        integral = p.pid_discount * integral + e;
        const double derivative = has_prev ? e - prev_error : 0.0;
        has_prev = true;
        prev_error = e;
        const double next = std::exp(std::log(omega) - (p.pid_kp * e + p.pid_ki * integral + p.pid_kd * derivative));
        const double balance = std::fabs(std::log10(ratio));
        if (balance < best_balance) {
            best_balance = balance;
            best_weight = next;
        }
        return next;
    }
};

// KKT measures of one candidate point, in the original space (solver.rs:133-152).
struct Kkt {
    double primal_obj = 0, dual_obj = 0, primal_res = 0, dual_res = 0;
    double gap() const { return std::fabs(primal_obj - dual_obj); }
    bool converged(double eps, double b_norm, double c_norm) const {
        return primal_res <= eps * (1.0 + b_norm) && dual_res <= eps * (1.0 + c_norm) &&
               gap() <= eps * (1.0 + std::fabs(primal_obj) + std::fabs(dual_obj));
    }
};

template <class B>
class Solver {
    using Vec = typename B::Vector;  // owning handle
    using Prob = typename B::Problem;

   public:
    // Solver::new (solver.rs:185-212): load, then thirteen vectors, zero except
    // c. The LP itself is not kept: only m, n, c0, the sense and the two
    // norms that run() needs.
    // With params.gpu_scaling, `scaled` is ignored: the backend uploads the LP
    // unscaled and scales it on the device, and scaled_ receives the factors.
    // With host scaling, `relabeled` (when not empty) is the column order the
    // caller already applied to `scaled`.
    Solver(B& b, const LpView& problem, const sp_params& params, ScaledProblem scaled, ColumnOrder relabeled = {})
        : b_(b), p_(params), scaled_(std::move(scaled)), cols_(std::move(relabeled)), m_(problem.m), n_(problem.n),
          nnz_(problem.nnz), c0_(problem.c0), maximize_(problem.maximize) {
        if (params.hot_table < -1 || params.hot_table > 1) throw Error(SP_ERR_INVALID, "hot_table must be -1, 0 or 1");
        if (params.reorder_rows < -1 || params.reorder_rows > 1)
            throw Error(SP_ERR_INVALID, "reorder_rows must be -1, 0 or 1");
        b_norm_ = finite_bound_norm(problem.lc, problem.uc, problem.m);
        c_norm_ = norm2(problem.c, problem.n);
        // Columns relabeled by popularity (the solution is returned in the
        // original order; see ColumnOrder).
        if (params.gpu_scaling) {
            // The relabeling (when asked for) is computed and applied on the device.
            dev_ = b_.load_scaled(problem, scaling_params_of(params), params.spmv_autotune != 0, scaled_,
                                  params.reorder_columns ? &cols_ : nullptr);
        } else {
            if (params.reorder_columns && cols_.empty()) {
                cols_ = popularity_order(problem);
                relabel(scaled_, cols_);
            }
            dev_ = b_.load(scaled_, params.spmv_autotune != 0);
        }
        // Column panels for the long rows of a wide A' (the gathered x̄ far
        // beyond an XCD's L2), with the rows relabeled long rows first; large
        // LPs get the same row relabeling without panels (reorder_rows), so
        // that each wave's gathers fall on fewer cache lines and the most
        // gathered entries of y come first, for the A'ᵀ side's hot table.
        const int shift = params.column_panels >= 0 ? params.column_panels : (n_ >= (int64_t(1) << 21) ? 16 : 0);
        if (shift || params.reorder_rows > 0 || (params.reorder_rows < 0 && nnz_ >= kReorderRowsMinNnz)) {
            rows_ = long_rows_first(problem, cols_, kPanelMinRow);
            b_.permute_rows(dev_.get(), rows_.order);
        }
        if (shift) b_.set_panels(dev_.get(), shift);
        b_.plan_hot(dev_.get(), params.hot_table);
        // The scaled matrices live on the device now; the host keeps the factors.
        scaled_.a = Csr{};
        scaled_.at = Csr{};
        const int64_t m = m_, n = n_;
        x_ = b_.zeros(n), y_ = b_.zeros(m), x_t_ = b_.zeros(n), y_t_ = b_.zeros(m), x_bar_ = b_.zeros(n);
        x_anchor_ = b_.zeros(n), y_anchor_ = b_.zeros(m), ax_ = b_.zeros(m), aty_ = b_.zeros(n);
        tmp_n_ = b_.zeros(n), tmp_m_ = b_.zeros(m), tmp_m2_ = b_.zeros(m);
        c_ = b_.upload(scaled_.c.data(), n);
        slots_ = b_.zeros(kSlotsAllocated);
        host_slots_ = b_.host_buffer(kSlotsAllocated);
        if (params.precision == 1) {
            if (!params.fused) throw Error(SP_ERR_INVALID, "mixed precision needs the fused path");
            b_.ensure_fp32(dev_.get());
            dx32_ = b_.zeros32(n), dy32_ = b_.zeros32(m), dx_bar32_ = b_.zeros32(n);
            ex32_ = b_.zeros32(n), ey32_ = b_.zeros32(m);
            atya_ = b_.zeros(n), axa_ = b_.zeros(m);
        }
    }

    // Solver::run (solver.rs:311-438). x_out/y_out receive the unscaled x_t, y_t.
    sp_solution run(double* x_out, double* y_out) {
        start();
        while (!advance(UINT64_MAX)) {
        }
        return finish(x_out, y_out);
    }

    // The part of run() before the loop (solver.rs:312-330): the clock starts,
    // power iteration sets η, and ω starts at 1 (or the norm ratio).
    void start() {
        power_begin();
        while (!power_done_) {
            const int rounds = power_rounds_next();
            for (int r = 0; r < rounds; ++r) {
                b_.spmv(dev_.get(), power_v_.get(), tmp_m_.get());
                b_.spmv_t(dev_.get(), tmp_m_.get(), power_v_.get());
                b_.dot_into(power_v_.get(), power_v_.get(), slot(kPower + r));
                b_.scale_inv_sqrt(power_v_.get(), slot(kPower + r));
            }
            b_.fetch_vec_async(slots_.get(), kPower + rounds, host_slots_.get());
            b_.sync();
            power_take(host_slots_.get() + kPower, rounds);
        }
        power_end();
    }

    // ---- start() in phases, for a caller that batches the power iteration
    // of many solvers (the serving loop) ----
    //
    // Largest singular value of A' by power iteration on A'ᵀA'
    // (solver.rs:215-242). power_begin() starts the clock and enqueues the
    // start vector's normalisation; then, in blocks of at most power_block()
    // rounds, the caller runs for r < power_rounds_next():
    //     tmp_m = A' v;  v = A'ᵀ tmp_m;  slots[slot_power() + r] = v·v;  v *= 1 / sqrt(that)
    // and passes the read-back ‖v‖² values to power_take(), until
    // power_done(); power_end() sets η, ω and the PID. The rounds after the
    // stopping one (inside a block) touch only v and tmp_m, which are not
    // used afterwards, so the result is the reference's.
    struct PowerWork {
        sp_problem* problem;
        sp_vec *v, *tmp_m;
        double* slots;  // device result slots
    };
    static constexpr int power_block() { return kPowerBlock; }
    static constexpr int slot_power() { return kPower; }
    bool power_done() const { return power_done_; }
    PowerWork power_work() { return PowerWork{dev_.get(), power_v_.get(), tmp_m_.get(), slots_->d}; }
    int power_rounds_next() const {
        return (int)std::min<int64_t>(kPowerBlock, p_.power_iterations - power_rounds_);
    }

    void power_begin() {
        start_ = clock::now();
        started_ = true;
        power_lambda_ = 0.0;
        power_rounds_ = 0;
        power_trivial_ = n_ == 0 || nnz_ == 0;
        power_done_ = power_trivial_ || p_.power_iterations <= 0;
        if (power_trivial_) return;
        std::vector<double> start((size_t)n_);
        for (int64_t j = 0; j < n_; ++j) {
            const int64_t orig = cols_.empty() ? j : cols_.order[(size_t)j];
            start[(size_t)j] = 1.0 + 0.01 * (double)(orig % 7);
        }
        power_v_ = b_.upload(start.data(), n_);
        // v *= 1 / ‖v‖, on the device (the same arithmetic as on the host).
        b_.dot_into(power_v_.get(), power_v_.get(), slot(kPowerStart));
        b_.scale_inv_sqrt(power_v_.get(), slot(kPowerStart));
    }

    // `norms` holds v·v of the `rounds` rounds just run.
    void power_take(const double* norms, int rounds) {
        bool done = false;
        for (int r = 0; r < rounds && !done; ++r) {
            const double norm = std::sqrt(norms[r]);
            if (std::isnan(norm) || norm <= 0.0 || std::isinf(norm))
                throw Error(SP_ERR_NUMERICAL, "numerical breakdown: power iteration collapsed");
            done = std::fabs(norm - power_lambda_) <= 1e-6 * norm;
            power_lambda_ = norm;
        }
        power_rounds_ += rounds;
        power_done_ = done || power_rounds_ >= p_.power_iterations;
    }

    void power_end() {
        // Power iteration underestimates; pad slightly so the step stays safe.
        const double norm = power_trivial_ ? 1.0 : std::sqrt(power_lambda_) * 1.01;
        power_v_.reset();
        eta_ = p_.step_size_factor / norm;
        omega_ = p_.bound_objective_rescaling ? 1.0 : (1.0 + c_norm_) / (1.0 + b_norm_);
        pid_ = PidState(omega_);
    }

    // Runs the loop (solver.rs:332-420) for at most `budget` more iterations;
    // returns true once it has terminated.
    bool advance(uint64_t budget) {
        const sp_params& p = p_;
        for (uint64_t done = 0; !terminated_ && done < budget; ++done) {
            const double tau = eta_ / omega_, sigma = eta_ * omega_;
            const uint64_t next_total = total_ + 1;
            const bool check = next_total % p.check_every == 0 || next_total >= p.max_iterations;
            if (p.fused && !check && k_ != 0) {
                // Nothing of this iteration is observed but x, y and x_bar.
                const double w = (double)(k_ + 1) / (double)(k_ + 2);
                if (mixed_now()) {
                    enter_plain();
                    b_.plain_iteration_mixed(dev_.get(), dx32_.get(), dy32_.get(), x_.get(), y_.get(), atya_.get(),
                                             axa_.get(), ex32_.get(), ey32_.get(), dx_bar32_.get(), tau, sigma, w,
                                             p.reflection);
                } else {
                    b_.plain_iteration(dev_.get(), x_.get(), y_.get(), x_anchor_.get(), y_anchor_.get(),
                                       x_bar_.get(), tau, sigma, w, p.reflection);
                    if (p.precision == 1) ++fp64_plain_;
                }
                total_ = next_total;
                k_ += 1;
                continue;
            }
            if (shadow_live_) {
                // Checks and restarts work on the FP64 iterates.
                b_.delta_to(dx32_.get(), x_.get(), x_.get());
                b_.delta_to(dy32_.get(), y_.get(), y_.get());
                shadow_live_ = false;
            }
            if (p.fused) {
                if (deferred_step(check, tau, sigma, next_total)) break;
                continue;
            } else {
                b_.spmv_t(dev_.get(), y_.get(), aty_.get());
                b_.primal_step(dev_.get(), x_.get(), aty_.get(), tau, x_t_.get(), x_bar_.get());
                b_.spmv(dev_.get(), x_bar_.get(), ax_.get());
                b_.dual_step(dev_.get(), y_.get(), ax_.get(), sigma, y_t_.get());
            }
            total_ = next_total;
            // At k = 0 the iterate is the anchor, so this measures r(z_anchor).
            const double residual = (check || k_ == 0) ? fixed_point_residual(eta_, omega_) : 0.0;
            if (k_ == 0) anchor_residual_ = residual;
            double step_dual = 0.0, step_obj = 0.0;
            if (check) {
                kkt_ = kkt();
                step_dual = std::sqrt(
                    b_.step_dual_residual_sq(dev_.get(), x_.get(), x_t_.get(), aty_.get(), tmp_n_.get(), tau));
                if (p.termination == 1) step_obj = step_dual_objective(tau);
            }
            const double w = (double)(k_ + 1) / (double)(k_ + 2);
            b_.halpern_step(x_.get(), x_t_.get(), x_anchor_.get(), w, p.reflection);
            b_.halpern_step(y_.get(), y_t_.get(), y_anchor_.get(), w, p.reflection);
            k_ += 1;
            if (!check) continue;
            checks_ += 1;
            log_check(residual);
            Kkt test = kkt_;
            if (p.termination == 1) {
                test.dual_obj = step_obj;
                test.dual_res = step_dual;
            }
            if (test.converged(p.tolerance, b_norm_, c_norm_)) {
                stop(SP_OPTIMAL);
                break;
            }
            watch_progress(test);
            const int inf = infeasibility();
            if (inf >= 0) {
                stop(inf);
                break;
            }
            if (total_ >= p.max_iterations) {
                stop(SP_ITERATION_LIMIT);
                break;
            }
            if (p.time_limit_s >= 0.0 && elapsed() >= p.time_limit_s) {
                stop(SP_TIME_LIMIT);
                break;
            }
            restart_decision(residual, step_dual, [&] {
                return std::make_pair(std::sqrt(b_.distance_sq(x_t_.get(), x_anchor_.get())),
                                      std::sqrt(b_.distance_sq(y_t_.get(), y_anchor_.get())));
            });
        }
        return terminated_;
    }

    uint64_t iterations() const { return total_; }
    // The loaded (scaled, relabeled) problem.
    const sp_problem* problem() const { return dev_.get(); }
    bool terminated() const { return terminated_; }

    // ---- for a caller that runs the plain iterations itself (the serving loop) ----

    // Plain iterations (no check, k != 0) that can run from here before the
    // next iteration that is not one; 0 when that next one is due now.
    uint64_t plain_run() const {
        if (terminated_ || !p_.fused || k_ == 0) return 0;
        const uint64_t ce = p_.check_every;
        uint64_t next = (total_ / ce + 1) * ce;
        next = std::min(next, std::max(p_.max_iterations, total_ + 1));
        return next - 1 - total_;
    }

    // Whether the next iteration is a check iteration of the fused path.
    bool check_due() const {
        if (terminated_ || !p_.fused) return false;
        const uint64_t next_total = total_ + 1;
        return next_total % p_.check_every == 0 || next_total >= p_.max_iterations;
    }

    // A check iteration in two halves, so that a caller can run many solvers'
    // checks behind one stream synchronization: check_begin() enqueues the
    // iteration and the readback of its results; after the caller has
    // synchronized the backend's stream, check_end() takes the decisions and
    // returns whether the solve has ended. Together they are advance(1) on a
    // check iteration.
    void check_begin() {
        if (!check_due()) throw Error(SP_ERR_INVALID, "check_begin: no check is due");
        if (shadow_live_) {
            b_.delta_to(dx32_.get(), x_.get(), x_.get());
            b_.delta_to(dy32_.get(), y_.get(), y_.get());
            shadow_live_ = false;
        }
        enqueue_step(true, eta_ / omega_, eta_ * omega_, total_ + 1);
        check_pending_ = true;
    }
    bool check_end() {
        if (!check_pending_) throw Error(SP_ERR_INVALID, "check_end: no check_begin before it");
        check_pending_ = false;
        return decide();
    }

    // ---- for a caller that batches non-plain iterations across solvers ----
    //
    // begin_step() describes the device work of the next iteration that is
    // not a plain one (a check iteration, or the first iteration of a restart
    // epoch) and does its bookkeeping; the caller then runs exactly the
    // operations of enqueue_step on these operands (together with other
    // solvers' work). For a check, the caller reads the result slots back and
    // passes them to end_step(), which takes the decisions. Restart copies are
    // appended to `copies` (source, destination) for the caller to run.
    struct StepWork {
        bool check = false, k_was_zero = false, step_implied = false;
        bool rebuild = false;  // mixed precision: x, y += displacement first
        int fix_base = 0;      // slot of the fixed-point sums
        sp_problem* problem = nullptr;
        sp_vec *x, *y, *x_t, *y_t, *x_bar, *x_anchor, *y_anchor, *ax, *aty, *tmp_n, *tmp_m, *tmp_m2, *c;
        // Check iterations: A'x_t, A'ᵀy_t, y_t - y_anchor, its A'ᵀ product,
        // x_t - x_anchor and its A' product (null otherwise).
        sp_vec *axt = nullptr, *atyt = nullptr, *dya = nullptr, *atdya = nullptr, *dxa = nullptr, *adxa = nullptr;
        sp_vec32 *dx = nullptr, *dy = nullptr;
        double* slots = nullptr;  // device, slot_count() values
        double tau = 0, sigma = 0, w = 0, rho = 0;
    };
    static constexpr int slot_count() { return kSlotCount; }
    static constexpr int slot_fix() { return kFix; }
    static constexpr int slot_primal() { return kPrimal; }
    static constexpr int slot_dual() { return kDual; }
    static constexpr int slot_step() { return kStep; }
    static constexpr int slot_step_rows() { return kStepRows; }
    static constexpr int slot_step_dots() { return kStepDots; }
    static constexpr int slot_dual_ray() { return kDualRay; }
    static constexpr int slot_primal_ray() { return kPrimalRay; }
    static constexpr int slot_dist_x() { return kDistX; }
    static constexpr int slot_dist_y() { return kDistY; }

    StepWork begin_step() {
        if (terminated_ || !p_.fused) throw Error(SP_ERR_INVALID, "begin_step: not running the fused path");
        const bool check = check_due();
        if (!check && k_ != 0) throw Error(SP_ERR_INVALID, "begin_step: a plain iteration is due");
        StepWork s;
        s.rebuild = shadow_live_;
        shadow_live_ = false;
        s.check = check;
        s.k_was_zero = k_ == 0;
        s.step_implied = p_.termination == 1;
        s.fix_base = check ? kFix : kAnchorFix;
        s.problem = dev_.get();
        s.x = x_.get(), s.y = y_.get(), s.x_t = x_t_.get(), s.y_t = y_t_.get(), s.x_bar = x_bar_.get();
        s.x_anchor = x_anchor_.get(), s.y_anchor = y_anchor_.get(), s.ax = ax_.get(), s.aty = aty_.get();
        s.tmp_n = tmp_n_.get(), s.tmp_m = tmp_m_.get(), s.tmp_m2 = tmp_m2_.get(), s.c = c_.get();
        if (check) {
            ensure_check_vectors();
            s.axt = ck_axt_.get(), s.atyt = ck_atyt_.get(), s.dya = ck_dya_.get(), s.atdya = ck_atdya_.get();
            s.dxa = ck_dxa_.get(), s.adxa = ck_adxa_.get();
        }
        s.dx = dx32_ ? dx32_.get() : nullptr;
        s.dy = dy32_ ? dy32_.get() : nullptr;
        s.slots = slots_->d;
        s.tau = eta_ / omega_;
        s.sigma = eta_ * omega_;
        s.w = (double)(k_ + 1) / (double)(k_ + 2);
        s.rho = p_.reflection;
        // The bookkeeping of enqueue_step.
        total_ += 1;
        if (check) {
            pending_tau_ = s.tau;
            pending_k_was_zero_ = s.k_was_zero;
            check_pending_ = true;
        } else {
            anchor_pending_ = true;
            anchor_omega_ = omega_;
        }
        k_ += 1;
        return s;
    }

    bool end_step(const double* slots, std::vector<std::pair<sp_vec*, sp_vec*>>* copies) {
        if (!check_pending_) throw Error(SP_ERR_INVALID, "end_step: no check is pending");
        check_pending_ = false;
        std::copy(slots, slots + kSlotCount, host_slots_.get());
        copy_sink_ = copies;
        try {
            const bool done = decide();
            copy_sink_ = nullptr;
            return done;
        } catch (...) {
            copy_sink_ = nullptr;
            throw;
        }
    }

    // What enter_plain() would enqueue, described for a batching caller: the
    // base's FP64 products, zero displacements and the anchors' offsets from
    // the base. Returns false when there is nothing to do.
    struct PlainEntry {
        sp_problem* problem = nullptr;
        sp_vec *x, *y, *x_anchor, *y_anchor, *atya, *axa;
        sp_vec32 *dx, *dy, *ex, *ey;
    };
    bool begin_plain(PlainEntry& e) {
        if (!mixed_now() || shadow_live_) return false;
        e.problem = dev_.get();
        e.x = x_.get(), e.y = y_.get(), e.x_anchor = x_anchor_.get(), e.y_anchor = y_anchor_.get();
        e.atya = atya_.get(), e.axa = axa_.get(), e.dx = dx32_.get(), e.dy = dy32_.get();
        e.ex = ex32_.get(), e.ey = ey32_.get();
        shadow_live_ = true;
        return true;
    }

    // Whether plain iterations run in mixed precision now.
    bool mixed_now() const { return p_.precision == 1 && !fp64_fallback_; }

    // Brings the iterates into the form plain iterations use: in mixed
    // precision, the current FP64 iterates become the base, with their FP64
    // products, zero FP32 displacements and the anchors' FP32 offsets from the
    // base.
    void enter_plain() {
        if (!mixed_now() || shadow_live_) return;
        b_.spmv_t(dev_.get(), y_.get(), atya_.get());
        b_.spmv(dev_.get(), x_.get(), axa_.get());
        b_.delta_from(x_.get(), x_.get(), dx32_.get());
        b_.delta_from(y_.get(), y_.get(), dy32_.get());
        b_.delta_from(x_anchor_.get(), x_.get(), ex32_.get());
        b_.delta_from(y_anchor_.get(), y_.get(), ey32_.get());
        shadow_live_ = true;
    }

    // The state plain iterations read and write, for a batch table.
    struct PlainState {
        sp_problem* problem;
        sp_vec *x, *y, *x_bar, *x_anchor, *y_anchor, *atya, *axa;
        sp_vec32 *dx, *dy, *dx_bar, *ex, *ey;
        double tau, sigma, rho;
        uint64_t k;
    };
    PlainState plain_state() {
        return PlainState{dev_.get(),
                          x_.get(),
                          y_.get(),
                          x_bar_.get(),
                          x_anchor_.get(),
                          y_anchor_.get(),
                          atya_ ? atya_.get() : nullptr,
                          axa_ ? axa_.get() : nullptr,
                          dx32_ ? dx32_.get() : nullptr,
                          dy32_ ? dy32_.get() : nullptr,
                          dx_bar32_ ? dx_bar32_.get() : nullptr,
                          ex32_ ? ex32_.get() : nullptr,
                          ey32_ ? ey32_.get() : nullptr,
                          eta_ / omega_,
                          eta_ * omega_,
                          p_.reflection,
                          k_};
    }

    // Average milliseconds of one launch of the plain iteration's main kernel
    // of `side` (0: A'ᵀ, 1: A'), launched `reps` times back to back on copies
    // of what it writes (measurement only: the solver's state is unchanged).
    double plain_kernel_time(int side, int reps) {
        if (!p_.fused) throw Error(SP_ERR_INVALID, "plain_kernel_time: needs the fused path");
        const bool mixed = mixed_now();
        if (mixed) enter_plain();
        typename B::PlainArgs a{};
        a.x = x_.get(), a.y = y_.get(), a.x_anchor = x_anchor_.get(), a.y_anchor = y_anchor_.get();
        a.x_bar = x_bar_.get();
        if (mixed) {
            a.dx = dx32_.get(), a.dy = dy32_.get(), a.dx_bar = dx_bar32_.get(), a.ex = ex32_.get();
            a.ey = ey32_.get(), a.atya = atya_.get(), a.axa = axa_.get();
        }
        a.tau = eta_ / omega_, a.sigma = eta_ * omega_, a.w = (double)(k_ + 1) / (double)(k_ + 2);
        a.rho = p_.reflection;
        return b_.plain_kernel_time(dev_.get(), side, mixed, a, reps);
    }

    // Accounts for `steps` plain iterations that the caller ran.
    void plain_done(uint64_t steps) {
        if (p_.precision == 1 && !mixed_now()) fp64_plain_ += steps;
        total_ += steps;
        k_ += steps;
    }

    // solver.rs:422-437. Before termination, status is -1 and the figures are
    // those of the last check.
    sp_solution finish(double* x_out, double* y_out) {
        if ((x_out || y_out) && n_ + m_ >= kDeviceUnscaleFrom) {
            // Unscaled on the device (scaling.rs:114-121), the relabelings
            // undone by the same pass; bitwise the host's arithmetic below.
            if (x_out) b_.unscale_to_host(x_t_.get(), scaled_.col_scale, scaled_.bound_scale, &cols_.order, false, x_out);
            if (y_out) b_.unscale_to_host(y_t_.get(), scaled_.row_scale, scaled_.obj_scale, &rows_.order, true, y_out);
        } else if (x_out || y_out) {
            std::vector<double> xs((size_t)n_), ys((size_t)m_);
            b_.download(x_t_.get(), xs.data(), n_);
            b_.download(y_t_.get(), ys.data(), m_);
            if (!rows_.empty()) {
                std::vector<double> yo((size_t)m_);
                for (int64_t i = 0; i < m_; ++i) yo[(size_t)rows_.order[(size_t)i]] = ys[(size_t)i];
                ys.swap(yo);
            }
            const std::vector<double> x = scaled_.unscale_primal(xs);
            const std::vector<double> y = scaled_.unscale_dual(ys);
            if (x_out && cols_.empty()) std::copy(x.begin(), x.end(), x_out);
            if (x_out && !cols_.empty())
                for (int64_t q = 0; q < n_; ++q) x_out[cols_.order[(size_t)q]] = x[(size_t)q];
            if (y_out) std::copy(y.begin(), y.end(), y_out);
        }
        const double sign = maximize_ ? -1.0 : 1.0;
        sp_solution s{};
        s.status = terminated_ ? status_ : -1;
        s.objective = sign * kkt_.primal_obj;
        s.dual_objective = sign * kkt_.dual_obj;
        s.primal_residual = kkt_.primal_res;
        s.dual_residual = kkt_.dual_res;
        s.gap = kkt_.gap();
        s.iterations = total_;
        s.restarts = restarts_;
        s.solve_time_s = terminated_ ? solve_time_ : elapsed();
        s.final_omega = omega_;
        s.mixed_fp64_iterations = fp64_plain_;
        return s;
    }

   private:
    // Result slots of the deferred check (fused path): everything a check
    // needs is reduced on the device and read back in one transfer.
    enum Slot {
        kFix = 0,         // dx², dy², cross of the fixed-point residual (3)
        kAnchorFix = 3,   // the same for a k = 0 iteration between checks (3)
        kPrimal = 6,      // primal_stats sums (2)
        kDual = 8,        // dual_stats sums (3)
        kStep = 11,       // step_dual_residual_sq (1)
        kStepRows = 12,   // dual_stats(y_t, c) sums, step-implied termination (3)
        kStepDots = 15,   // c·x_t, aty·x_t, x_t·x_t, x·x_t (4)
        kDualRay = 19,    // dual_ray_stats sums (5)
        kPrimalRay = 24,  // primal_ray_stats sums (4)
        kDistX = 28,      // ‖x_t - x_anchor‖² (1)
        kDistY = 29,      // ‖y_t - y_anchor‖² (1)
        kSlotCount = 30,
        kPower = 30,      // power iteration: one block of ‖v‖² (kPowerBlock)
        kPowerStart = 38, // power iteration: ‖v‖² of the start vector
        kSlotsAllocated = 39
    };
    static constexpr int kPowerBlock = 8;
    // From this many entries of x and y together, finish() unscales on the device.
    static constexpr int64_t kDeviceUnscaleFrom = int64_t(1) << 20;

    void enqueue_fixed_point(int base) {
        b_.sub(x_t_.get(), x_.get(), tmp_n_.get());
        b_.sub(y_t_.get(), y_.get(), tmp_m2_.get());
        b_.spmv(dev_.get(), tmp_n_.get(), tmp_m_.get());
        b_.dot_into(tmp_n_.get(), tmp_n_.get(), slot(base));
        b_.dot_into(tmp_m2_.get(), tmp_m2_.get(), slot(base + 1));
        b_.dot_into(tmp_m2_.get(), tmp_m_.get(), slot(base + 2));
    }

    // fixed_point_residual's closing arithmetic (solver.rs:283-287).
    static double residual_of(const double* r, double eta, double omega) {
        const double r2 = omega / eta * r[0] + r[1] / (eta * omega) + 2.0 * r[2];
        if (!std::isfinite(r2)) throw Error(SP_ERR_NUMERICAL, "numerical breakdown: iterates are not finite");
        return std::sqrt(r2 > 0.0 ? r2 : 0.0);
    }

    // One non-plain iteration of the fused path: the reference's calls in
    // the reference's order, with the reductions written to result slots and
    // the infeasibility rays and restart distances formed speculatively, so
    // that a check costs one device-to-host transfer. The decisions that
    // follow are those of solver.rs:340-419, on the same values. Returns true
    // when the loop has ended.
    bool deferred_step(bool check, double tau, double sigma, uint64_t next_total) {
        enqueue_step(check, tau, sigma, next_total);
        if (!check) return false;
        b_.sync();
        return decide();
    }

    // The device work of deferred_step; on a check it ends with the
    // asynchronous readback of the result slots into host_slots_.
    void enqueue_step(bool check, double tau, double sigma, uint64_t next_total) {
        const sp_params& p = p_;
        b_.spmv_t_primal(dev_.get(), y_.get(), x_.get(), tau, aty_.get(), x_t_.get(), x_bar_.get());
        b_.spmv_dual(dev_.get(), x_bar_.get(), y_.get(), sigma, ax_.get(), y_t_.get());
        total_ = next_total;
        const bool k_was_zero = k_ == 0;
        const double w = (double)(k_ + 1) / (double)(k_ + 2);
        if (check) {
            enqueue_check(tau, w);
        } else {
            if (k_was_zero) {
                // r(z_anchor), needed at the next check (solver.rs:341-346).
                enqueue_fixed_point(kAnchorFix);
                anchor_pending_ = true;
                anchor_omega_ = omega_;
            }
            b_.halpern_step(x_.get(), x_t_.get(), x_anchor_.get(), w, p.reflection);
            b_.halpern_step(y_.get(), y_t_.get(), y_anchor_.get(), w, p.reflection);
        }
        k_ += 1;
        if (!check) return;
        b_.fetch_vec_async(slots_.get(), kSlotCount, host_slots_.get());
        pending_tau_ = tau;
        pending_k_was_zero_ = k_was_zero;
    }

    // The rest of a check iteration after its two fused SpMVs: the
    // fixed-point residual's sums (solver.rs:279-287), the KKT sums (kkt,
    // step_dual_residual_sq and, for step-implied termination, the dual
    // objective's terms), the speculative infeasibility rays (solver.rs:291-309)
    // and restart distances (solver.rs:403-404), each into its result slot,
    // then the Halpern updates. Every reduction reads x before its Halpern
    // update and nothing the update writes, so the differences run as one
    // launch, the five products into vectors of their own, and all the
    // reductions as one launch (plus its finishing pass), each bitwise the
    // method it stands for (Hip::upload_ops).
    void enqueue_check(double tau, double w) {
        if (!check_ops_) build_check_ops();
        b_.run_subs(check_ops_.get());
        b_.spmv(dev_.get(), tmp_n_.get(), tmp_m_.get());
        b_.spmv(dev_.get(), x_t_.get(), ck_axt_.get());
        b_.spmv_t(dev_.get(), y_t_.get(), ck_atyt_.get());
        b_.spmv_t(dev_.get(), ck_dya_.get(), ck_atdya_.get());
        b_.spmv(dev_.get(), ck_dxa_.get(), ck_adxa_.get());
        b_.run_reductions(check_ops_.get(), 1.0 / tau);  // 1/τ as step_dual_residual_sq forms it
        b_.run_halperns(check_ops_.get(), w, p_.reflection);
    }

    // The vectors only check iterations of the fused path use (made at the
    // first one).
    void ensure_check_vectors() {
        if (ck_axt_) return;
        const int64_t m = m_, n = n_;
        ck_axt_ = b_.zeros(m), ck_atyt_ = b_.zeros(n), ck_dya_ = b_.zeros(m), ck_atdya_ = b_.zeros(n);
        ck_dxa_ = b_.zeros(n), ck_adxa_ = b_.zeros(m);
    }

    // The operand table of enqueue_check.
    void build_check_ops() {
        const int64_t m = m_, n = n_;
        ensure_check_vectors();
        const sp_problem* P = dev_.get();
        OpList l;
        auto sub = [&](const Vec& a, const Vec& bv, Vec& out) {
            COp o;
            o.a0 = a->d, o.a1 = bv->d, o.o0 = out->d, o.n = out->len;
            l.subs.push_back(o);
        };
        sub(x_t_, x_, tmp_n_);           // x_t - x        (fixed point)
        sub(y_t_, y_, tmp_m2_);          // y_t - y        (fixed point)
        sub(y_t_, y_anchor_, ck_dya_);   // y_t - y_anchor (dual ray)
        sub(x_t_, x_anchor_, ck_dxa_);   // x_t - x_anchor (primal ray)
        // Operands in the order of the kind's reduction functor in kernels.hip.
        auto red = [&](RedKind kind, std::vector<const double*> a, bool rows_and_cols, int64_t len, int at) {
            COp o;
            const double** f[] = {&o.a0, &o.a1, &o.a2, &o.a3, &o.a4, &o.a5, &o.a6, &o.a7};
            for (size_t i = 0; i < a.size(); ++i) *f[i] = a[i];
            o.kind = kind;
            o.m = rows_and_cols ? m : 0;
            o.n = rows_and_cols ? n : len;
            o.dst = slot(at);
            l.reds.push_back(o);
        };
        red(kRedDot, {tmp_n_->d, tmp_n_->d}, false, n, kFix);
        red(kRedDot, {tmp_m2_->d, tmp_m2_->d}, false, m, kFix + 1);
        red(kRedDot, {tmp_m2_->d, tmp_m_->d}, false, m, kFix + 2);
        red(kRedPrimal, {ck_axt_->d, P->lc, P->uc, P->pf, P->c, x_t_->d}, true, 0, kPrimal);
        red(kRedDual, {y_t_->d, P->lc, P->uc, P->c, ck_atyt_->d, P->lv, P->uv, P->df}, true, 0, kDual);
        red(kRedStep, {x_->d, x_t_->d, aty_->d, ck_atyt_->d, P->df}, false, n, kStep);
        l.reds.back().s0 = 1.0 / (eta_ / omega_);  // rewritten by run_reductions when τ changes
        if (p_.termination == 1) {
            red(kRedDual, {y_t_->d, P->lc, P->uc, P->c, c_->d, P->lv, P->uv, P->df}, true, 0, kStepRows);
            red(kRedDot, {c_->d, x_t_->d}, false, n, kStepDots);
            red(kRedDot, {aty_->d, x_t_->d}, false, n, kStepDots + 1);
            red(kRedDot, {x_t_->d, x_t_->d}, false, n, kStepDots + 2);
            red(kRedDot, {x_->d, x_t_->d}, false, n, kStepDots + 3);
        }
        red(kRedDualRay, {ck_dya_->d, P->lc, P->uc, ck_atdya_->d, P->lv, P->uv}, true, 0, kDualRay);
        red(kRedPrimalRay, {ck_adxa_->d, P->lc, P->uc, ck_dxa_->d, P->c, P->lv, P->uv}, true, 0, kPrimalRay);
        red(kRedDist, {x_t_->d, x_anchor_->d}, false, n, kDistX);
        red(kRedDist, {y_t_->d, y_anchor_->d}, false, m, kDistY);
        auto halpern = [&](Vec& z, const Vec& z_t, const Vec& anchor) {
            COp o;
            o.o0 = z->d, o.a0 = z_t->d, o.a1 = anchor->d, o.n = z->len;
            l.halperns.push_back(o);
        };
        halpern(x_, x_t_, x_anchor_);
        halpern(y_, y_t_, y_anchor_);
        check_ops_ = b_.upload_ops(l);
    }

    // The host decisions of a check iteration on the slots read back by
    // enqueue_step (after the stream has been synchronized).
    bool decide() {
        const sp_params& p = p_;
        const double* r = host_slots_.get();
        const double tau = pending_tau_;
        const bool k_was_zero = pending_k_was_zero_;
        if (anchor_pending_) {
            anchor_residual_ = residual_of(r + kAnchorFix, eta_, anchor_omega_);
            anchor_pending_ = false;
        }
        const double residual = residual_of(r + kFix, eta_, omega_);
        if (k_was_zero) anchor_residual_ = residual;
        const double f = scaled_.objective_factor();
        const sp_primal_stats_t ps = B::primal_from(r + kPrimal);
        const sp_dual_stats_t ds = B::dual_from(r + kDual);
        kkt_.primal_obj = ps.objective * f + c0_;
        kkt_.dual_obj = ds.objective * f + c0_;
        kkt_.primal_res = std::sqrt(ps.residual_sq);
        kkt_.dual_res = std::sqrt(ds.residual_sq);
        const double step_dual = std::sqrt(r[kStep]);
        checks_ += 1;
        log_check(residual);
        Kkt test = kkt_;
        if (p.termination == 1) {
            const double rows = B::dual_from(r + kStepRows).objective;
            const double* d = r + kStepDots;
            const double lx = d[0] - d[1] + (d[2] - d[3]) / tau;
            test.dual_obj = (rows + lx) * f + c0_;
            test.dual_res = step_dual;
        }
        if (test.converged(p.tolerance, b_norm_, c_norm_)) {
            stop(SP_OPTIMAL);
            return true;
        }
        watch_progress(test);
        const int inf = infeasibility_of(B::dual_ray_from(r + kDualRay), B::primal_ray_from(r + kPrimalRay));
        if (inf >= 0) {
            stop(inf);
            return true;
        }
        if (total_ >= p.max_iterations) {
            stop(SP_ITERATION_LIMIT);
            return true;
        }
        if (p.time_limit_s >= 0.0 && elapsed() >= p.time_limit_s) {
            stop(SP_TIME_LIMIT);
            return true;
        }
        restart_decision(residual, step_dual,
                         [&] { return std::make_pair(std::sqrt(r[kDistX]), std::sqrt(r[kDistY])); });
        return false;
    }

    // The progress line of solver.rs:367-378.
    void log_check(double residual) {
        if (p_.log_every_checks > 0 && checks_ % (uint64_t)p_.log_every_checks == 0) {
            std::fprintf(stderr, "%9llu %14.6e %14.6e %9.2e %9.2e %9.2e %9.2e %5llu %8.2fs %7llu %9.2e\n",
                         (unsigned long long)total_, kkt_.primal_obj, kkt_.dual_obj, kkt_.primal_res, kkt_.dual_res,
                         kkt_.gap(), omega_, (unsigned long long)restarts_, elapsed(), (unsigned long long)k_,
                         residual / anchor_residual_);
        }
    }

    // Restart decision on the fixed-point residual (solver.rs:396-419).
    // `distances` yields (‖x_t - x_anchor‖, ‖y_t - y_anchor‖).
    template <class D>
    void restart_decision(double residual, double step_dual, D&& distances) {
        const sp_params& p = p_;
        const double r0 = anchor_residual_;
        const bool restart = residual <= p.restart_sufficient * r0 ||
                             (residual <= p.restart_necessary * r0 && residual > prev_residual_) ||
                             (double)k_ >= p.restart_artificial * (double)total_;
        prev_residual_ = residual;
        if (!restart) return;
        const auto [dx, dy] = distances();
        const double rel_primal = kkt_.primal_res / (1.0 + b_norm_);
        // The PID balances against the step-based dual residual.
        const double rel_dual = step_dual / (1.0 + c_norm_);
        omega_ = pid_.update(p, omega_, dx, dy, rel_primal, rel_dual);
        if (copy_sink_) {
            copy_sink_->emplace_back(x_t_.get(), x_anchor_.get());
            copy_sink_->emplace_back(y_t_.get(), y_anchor_.get());
            copy_sink_->emplace_back(x_t_.get(), x_.get());
            copy_sink_->emplace_back(y_t_.get(), y_.get());
        } else {
            b_.copy(x_t_.get(), x_anchor_.get());
            b_.copy(y_t_.get(), y_anchor_.get());
            b_.copy(x_t_.get(), x_.get());
            b_.copy(y_t_.get(), y_.get());
        }
        k_ = 0;
        restarts_ += 1;
        prev_residual_ = INFINITY;
    }

    // The certificate tests of solver.rs:298 and :305.
    bool dual_ray_certifies(const sp_ray_stats_t& d) const {
        const double eps = p_.infeasibility_tolerance;
        return d.objective > 0.0 && std::sqrt(d.violation_sq) <= eps * d.objective &&
               d.objective > eps * std::sqrt(d.norm_sq);
    }
    bool primal_ray_certifies(const sp_ray_stats_t& pr) const {
        const double eps = p_.infeasibility_tolerance;
        return pr.objective < 0.0 && std::sqrt(pr.violation_sq) <= -eps * pr.objective &&
               -pr.objective > eps * std::sqrt(pr.norm_sq);
    }

    // infeasibility() on rays that were already reduced.
    int infeasibility_of(const sp_ray_stats_t& d, const sp_ray_stats_t& pr) {
        if (dual_ray_certifies(d)) return SP_PRIMAL_INFEASIBLE;
        near_certificate(d.objective, d.violation_sq, d.norm_sq);
        if (primal_ray_certifies(pr)) return SP_DUAL_INFEASIBLE;
        near_certificate(-pr.objective, pr.violation_sq, pr.norm_sq);
        return -1;
    }

    bool anchor_pending_ = false;
    double anchor_omega_ = 1.0;
    Vec slots_;  // this solver's result slots (device), read back by one transfer per check
    typename B::HostBuffer host_slots_;  // their pinned host copy
    double pending_tau_ = 0.0;
    bool pending_k_was_zero_ = false;
    bool check_pending_ = false;
    std::vector<std::pair<sp_vec*, sp_vec*>>* copy_sink_ = nullptr;  // restart copies deferred to a batching caller
    // Power iteration state (start() or its phases).
    Vec power_v_;
    double power_lambda_ = 0.0;
    int64_t power_rounds_ = 0;
    bool power_done_ = false, power_trivial_ = false;
    double* slot(int i) { return slots_->d + i; }

   private:
    using clock = std::chrono::steady_clock;

    double elapsed() const { return started_ ? std::chrono::duration<double>(clock::now() - start_).count() : 0.0; }

    void stop(int status) {
        status_ = status;
        terminated_ = true;
        solve_time_ = elapsed();
    }

    // solver.rs:245-258. Leaves A'ᵀy_t in tmp_n.
    Kkt kkt() {
        const double f = scaled_.objective_factor();
        b_.spmv(dev_.get(), x_t_.get(), tmp_m_.get());
        const sp_primal_stats_t ps = b_.primal_stats(dev_.get(), x_t_.get(), tmp_m_.get());
        b_.spmv_t(dev_.get(), y_t_.get(), tmp_n_.get());
        const sp_dual_stats_t ds = b_.dual_stats(dev_.get(), y_t_.get(), tmp_n_.get());
        Kkt k;
        k.primal_obj = ps.objective * f + c0_;
        k.dual_obj = ds.objective * f + c0_;
        k.primal_res = std::sqrt(ps.residual_sq);
        k.dual_res = std::sqrt(ds.residual_sq);
        return k;
    }

    // cuPDLPx's dual objective of y_t (solver.rs:263-271).
    double step_dual_objective(double tau) {
        const double rows = b_.dual_stats(dev_.get(), y_t_.get(), c_.get()).objective;
        const double lx = b_.dot(c_.get(), x_t_.get()) - b_.dot(aty_.get(), x_t_.get()) +
                          (b_.dot(x_t_.get(), x_t_.get()) - b_.dot(x_.get(), x_t_.get())) / tau;
        return (rows + lx) * scaled_.objective_factor() + c0_;
    }

    // ‖z - T(z)‖_P in the PDHG metric (solver.rs:275-288).
    double fixed_point_residual(double eta, double omega) {
        b_.sub(x_t_.get(), x_.get(), tmp_n_.get());
        b_.sub(y_t_.get(), y_.get(), tmp_m2_.get());
        b_.spmv(dev_.get(), tmp_n_.get(), tmp_m_.get());
        const double dx2 = b_.norm_sq(tmp_n_.get());
        const double dy2 = b_.norm_sq(tmp_m2_.get());
        const double cross = b_.dot(tmp_m2_.get(), tmp_m_.get());
        const double r2 = omega / eta * dx2 + dy2 / (eta * omega) + 2.0 * cross;
        if (!std::isfinite(r2)) throw Error(SP_ERR_NUMERICAL, "numerical breakdown: iterates are not finite");
        return std::sqrt(r2 > 0.0 ? r2 : 0.0);
    }

    // z_t - z_anchor as an infeasibility certificate (solver.rs:291-309); -1 if none.
    int infeasibility() {
        b_.sub(y_t_.get(), y_anchor_.get(), tmp_m_.get());
        b_.spmv_t(dev_.get(), tmp_m_.get(), tmp_n_.get());
        const sp_ray_stats_t d = b_.dual_ray_stats(dev_.get(), tmp_m_.get(), tmp_n_.get());
        if (dual_ray_certifies(d)) return SP_PRIMAL_INFEASIBLE;
        near_certificate(d.objective, d.violation_sq, d.norm_sq);
        b_.sub(x_t_.get(), x_anchor_.get(), tmp_n_.get());
        b_.spmv(dev_.get(), tmp_n_.get(), tmp_m_.get());
        const sp_ray_stats_t pr = b_.primal_ray_stats(dev_.get(), tmp_n_.get(), tmp_m_.get());
        if (primal_ray_certifies(pr)) return SP_DUAL_INFEASIBLE;
        near_certificate(-pr.objective, pr.violation_sq, pr.norm_sq);
        return -1;
    }

    // Mixed precision only: a ray that satisfies the certificate test with a
    // looser violation bound (kNearCertificate instead of the infeasibility
    // tolerance) cannot be completed with FP32 displacements, whose rounding
    // is far above 1e-8 relative; the rest of the solve then runs in FP64.
    static constexpr double kNearCertificate = 1e-4;
    void near_certificate(double objective, double violation_sq, double norm_sq) {
        if (p_.precision != 1 || fp64_fallback_) return;
        if (objective > 0.0 && std::sqrt(violation_sq) <= kNearCertificate * objective &&
            objective > p_.infeasibility_tolerance * std::sqrt(norm_sq))
            fp64_fallback_ = true;
    }

    // Mixed precision only, with params.mixed_stall_checks = N > 0: FP64 takes
    // over when the relative KKT error has not halved its best for N checks in
    // a row. Off by default: with the displacement re-based at every check the
    // FP32 rounding stays far below the progress of a check interval, and a
    // slow LP stalls by this measure in FP64 just as often.
    void watch_progress(const Kkt& k) {
        if (p_.precision != 1 || fp64_fallback_ || p_.mixed_stall_checks <= 0) return;
        const double err = std::max({k.primal_res / (1.0 + b_norm_), k.dual_res / (1.0 + c_norm_),
                                     k.gap() / (1.0 + std::fabs(k.primal_obj) + std::fabs(k.dual_obj))});
        if (err < 0.5 * best_err_) {
            best_err_ = err;
            stalled_checks_ = 0;
        } else if (++stalled_checks_ >= (uint64_t)p_.mixed_stall_checks) {
            fp64_fallback_ = true;
        }
    }

    B& b_;
    const sp_params p_;
    ScaledProblem scaled_;
    ColumnOrder cols_;  // empty unless params.reorder_columns
    RowOrder rows_;     // empty unless the rows are relabeled (column panels or reorder_rows)
    int64_t m_, n_, nnz_;
    double c0_;
    bool maximize_;
    double b_norm_ = 0.0, c_norm_ = 0.0;
    // Loop state of run() (solver.rs:316-330).
    clock::time_point start_{};
    bool started_ = false, terminated_ = false;
    int status_ = -1;
    double solve_time_ = 0.0;
    double eta_ = 1.0, omega_ = 1.0;
    PidState pid_{1.0};
    uint64_t total_ = 0, k_ = 0, restarts_ = 0, checks_ = 0;
    double anchor_residual_ = INFINITY, prev_residual_ = INFINITY;
    Kkt kkt_;
    Prob dev_;
    Vec x_, y_, x_t_, y_t_, x_bar_, x_anchor_, y_anchor_, ax_, aty_, tmp_n_, tmp_m_, tmp_m2_, c_;
    // Check iterations of the fused path (build_check_ops): A'x_t, A'ᵀy_t,
    // y_t - y_anchor and its A'ᵀ product, x_t - x_anchor and its A' product,
    // and the operand table over them.
    Vec ck_axt_, ck_atyt_, ck_dya_, ck_atdya_, ck_dxa_, ck_adxa_;
    typename B::Ops check_ops_;
    // Mixed precision: between checks x = x_ + dx32_, y = y_ + dy32_ (x_, y_
    // are the iterates at the last check, the base), with the base's FP64
    // products A'ᵀy_, A'x_ and the anchors' offsets from the base.
    typename B::Vector32 dx32_, dy32_, dx_bar32_, ex32_, ey32_;
    Vec atya_, axa_;
    bool shadow_live_ = false;
    bool fp64_fallback_ = false;
    double best_err_ = INFINITY;
    uint64_t stalled_checks_ = 0;
    uint64_t fp64_plain_ = 0;  // plain iterations that ran in FP64 in mixed mode
};

}  // namespace sp
