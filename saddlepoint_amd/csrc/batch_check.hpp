// Batched non-plain iterations for the serving loop.
//
// A check iteration of the fused driver (Solver::enqueue_step) is a fixed
// sequence of SpMVs, element-wise passes and reductions; so is the first
// iteration of a restart epoch, and the mixed-precision entry into plain
// iterations (Solver::enter_plain). Here that sequence runs once for a whole
// group of LPs: each operation becomes one launch (two for an SpMV with long
// rows) whose table holds every LP's operands, and all the reductions of a
// step are one launch (two when an entry spans more than one workgroup), as
// in Solver::enqueue_check. The kernels (sp_c* in kernels.hip) give each LP
// the partition and summation order of its single-LP launch, so every LP's
// results are bitwise those of Solver::advance.
//
// All tables of one phase are packed into one blob, copied to the device in
// one transfer from a pinned arena, then the launches run in sequence order.
#pragma once

#include <cstring>
#include <utility>
#include <vector>

#include "driver.hpp"
#include "solver.hpp"

namespace sp {

// Host mirrors of BDesc / BLong in kernels.hip (same layout); COp and CBlk
// are in backend.hpp.
struct BDesc {
    Desc d;
    int inst, pad0, pad1, pad2;
};
struct BLong {
    int inst, row, pbase, nch;
};
static_assert(sizeof(BDesc) == 32, "BDesc layout");
static_assert(sizeof(BLong) == 16, "BLong layout");

// Pinned host arena + device arena with bump allocation; reset once the
// stream has drained everything that read from them.
class TableArena {
   public:
    explicit TableArena(sp_backend* b) : b_(b) {}
    ~TableArena() {
        if (h_) (void)hipHostFree(h_);
        if (d_) (void)hipFree(d_);
    }
    void reset() { used_ = 0; }
    // Copies `bytes` of `src` to the device (stream-ordered); returns the
    // device address.
    char* upload(const void* src, size_t bytes) {
        if (bytes == 0) return d_;
        const size_t need = (used_ + bytes + 255) & ~size_t(255);
        if (need > cap_) {
            // Everything uploaded from the old arena has been consumed once the
            // stream drains.
            hip_check(hipStreamSynchronize(b_->stream), "hipStreamSynchronize");
            if (h_) (void)hipHostFree(h_);
            if (d_) (void)hipFree(d_);
            h_ = nullptr;
            d_ = nullptr;
            cap_ = std::max<size_t>(2 * cap_, std::max<size_t>(bytes * 2, size_t(1) << 20));
            hip_check(hipHostMalloc(reinterpret_cast<void**>(&h_), cap_, hipHostMallocDefault), "hipHostMalloc");
            hip_check(hipMalloc(reinterpret_cast<void**>(&d_), cap_), "hipMalloc");
            used_ = 0;
        }
        std::memcpy(h_ + used_, src, bytes);
        hip_check(hipMemcpyAsync(d_ + used_, h_ + used_, bytes, hipMemcpyHostToDevice, b_->stream),
                  "hipMemcpyAsync H2D");
        char* out = d_ + used_;
        used_ = (used_ + bytes + 255) & ~size_t(255);
        return out;
    }

   private:
    sp_backend* b_;
    char* h_ = nullptr;
    char* d_ = nullptr;
    size_t cap_ = 0, used_ = 0;
};

// One phase: tables packed into a blob, launches recorded against offsets.
class Phase {
   public:
    explicit Phase(sp_backend* b) : b_(b) {}

    // An SpMV over side A' (a = true) or A'ᵀ of each entry's problem, with
    // the epilogue of kernel k (finishing kernel fin for multi-chunk rows).
    // The matrix fields of `ops` are filled in here.
    void spmv(hipFunction_t k, hipFunction_t fin, std::vector<COp> ops, const std::vector<sp_problem*>& probs,
              bool a) {
        std::vector<BDesc> desc;
        std::vector<BLong> longs;
        for (size_t i = 0; i < ops.size(); ++i) {
            const DevCsr& M = a ? probs[i]->a : probs[i]->at;
            ops[i].rp = M.rp;
            ops[i].ci = M.ci;
            ops[i].va = M.va;
            ops[i].part = M.partials;
            for (const Desc& d : M.host_desc) desc.push_back(BDesc{d, (int)i, 0, 0, 0});
            for (const LongRow& l : M.host_longs) longs.push_back(BLong{(int)i, l.row, l.pbase, l.nch});
        }
        if (desc.empty()) return;
        const size_t o = put(ops), dd = put(desc);
        ls_.push_back(L{k, kSpmv, (unsigned)desc.size(), dd, o, 0});
        if (!longs.empty()) {
            const size_t ll = put(longs);
            ls_.push_back(L{fin, kFin, (unsigned)((longs.size() + 3) / 4), ll, o, (int)longs.size()});
        }
    }

    // An element-wise pass over each entry's [0, n).
    void elementwise(hipFunction_t k, const std::vector<COp>& ops) {
        std::vector<CBlk> blk;
        for (size_t i = 0; i < ops.size(); ++i)
            for (long long t = 0; t * kCTile < ops[i].n; ++t) blk.push_back(CBlk{(int)i, (int)t});
        if (blk.empty()) return;
        const size_t o = put(ops), bb = put(blk);
        ls_.push_back(L{k, kElem, (unsigned)blk.size(), bb, o, 0});
    }

    // Reductions of several kinds (each entry's kind set; sp_cred_any_f64),
    // every entry with the workgroup count of its single-LP launch: one
    // launch, and one finishing pass for the entries of more than one
    // workgroup.
    void reduce_any(std::vector<COp> ops, const Hip& h) {
        std::vector<CBlk> blk;
        std::vector<int> which;
        int pb = 0;
        for (size_t i = 0; i < ops.size(); ++i) {
            ops[i].K = red_sums(ops[i].kind);
            ops[i].G = h.reduce_grid(red_len(ops[i]));
            ops[i].pbase = pb;
            pb += ops[i].G * ops[i].K;
            for (int t = 0; t < ops[i].G; ++t) blk.push_back(CBlk{(int)i, t});
            if (ops[i].G > 1) which.push_back((int)i);
        }
        if (blk.empty()) return;
        partials_ = std::max(partials_, (size_t)pb);
        const size_t o = put(ops), bb = put(blk);
        ls_.push_back(L{b_->k_cred_any, kRed, (unsigned)blk.size(), bb, o, 0});
        if (!which.empty()) {
            const size_t ww = put(which);
            ls_.push_back(L{nullptr, kRedFin, (unsigned)which.size(), ww, o, 0});
        }
    }

    // Uploads the tables and enqueues the launches in order.
    void run(TableArena& arena, double*& partials, size_t& partials_cap) {
        if (ls_.empty()) return;
        if (partials_ > partials_cap) {
            hip_check(hipStreamSynchronize(b_->stream), "hipStreamSynchronize");
            if (partials) (void)hipFree(partials);
            partials = nullptr;
            partials_cap = std::max(partials_, 2 * partials_cap);
            hip_check(hipMalloc(reinterpret_cast<void**>(&partials), partials_cap * sizeof(double)), "hipMalloc");
        }
        char* base = arena.upload(blob_.data(), blob_.size());
        for (const L& l : ls_) {
            void* list = base + l.list;
            void* ops = base + l.ops;
            switch (l.type) {
                case kSpmv:
                    go(l.k, l.grid, list, ops);
                    break;
                case kFin:
                    go(l.k, l.grid, list, l.count, ops);
                    break;
                case kElem:
                    go(l.k, l.grid, list, ops);
                    break;
                case kRed:
                    go(l.k, l.grid, list, ops, partials);
                    break;
                case kRedFin:
                    go(b_->k_cred_finish_any, l.grid, list, ops, (const double*)partials);
                    break;
            }
        }
    }

   private:
    enum { kSpmv, kFin, kElem, kRed, kRedFin };
    struct L {
        hipFunction_t k;
        int type;
        unsigned grid;
        size_t list, ops;
        int count;
    };

    template <class T>
    size_t put(const std::vector<T>& v) {
        const size_t off = (blob_.size() + 63) & ~size_t(63);
        blob_.resize(off + v.size() * sizeof(T));
        if (!v.empty()) std::memcpy(blob_.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }

    template <class... Args>
    void go(hipFunction_t f, unsigned grid, Args... args) {
        void* params[] = {static_cast<void*>(&args)...};
        hip_check(hipModuleLaunchKernel(f, grid, 1, 1, 256, 1, 1, 0, b_->stream, params, nullptr),
                  "hipModuleLaunchKernel");
    }

    sp_backend* b_;
    std::vector<char> blob_;
    std::vector<L> ls_;
    size_t partials_ = 0;
};

using StepWork = Solver<DriverBackend>::StepWork;
using PlainEntry = Solver<DriverBackend>::PlainEntry;

// The Halpern steps of every entry (after its reductions, which read x and
// y before the update).
inline void enqueue_halpern(Phase& P, sp_backend* b, const std::vector<StepWork>& ws) {
    std::vector<COp> v;
    for (const StepWork& s : ws) {
        COp x, y;
        x.o0 = s.x->d, x.a0 = s.x_t->d, x.a1 = s.x_anchor->d, x.s0 = s.w, x.s1 = s.rho, x.n = s.x->len;
        y.o0 = s.y->d, y.a0 = s.y_t->d, y.a1 = s.y_anchor->d, y.s0 = s.w, y.s1 = s.rho, y.n = s.y->len;
        v.push_back(x);
        v.push_back(y);
    }
    P.elementwise(b->k_chalpern, v);
}

// The device work of Solver::enqueue_step for every entry of `ws` (all check
// iterations, or all first iterations of a restart epoch), in its order.
inline void enqueue_step_work(Phase& P, const Hip& h, const std::vector<StepWork>& ws) {
    if (ws.empty()) return;
    sp_backend* b = h.b;
    const bool check = ws[0].check;
    std::vector<sp_problem*> probs;
    for (const StepWork& s : ws) probs.push_back(s.problem);
    auto each = [&](auto f) {
        std::vector<COp> v;
        for (const StepWork& s : ws) f(v, s);
        return v;
    };
    auto slot = [](const StepWork& s, int i) { return s.slots + i; };

    // Mixed precision: the FP64 iterates, base + displacement.
    P.elementwise(b->k_cdelta_to, each([](std::vector<COp>& v, const StepWork& s) {
        if (!s.rebuild) return;
        COp x, y;
        x.o0 = s.x->d, x.a0 = s.x->d, x.f0 = s.dx->d, x.n = s.x->len;
        y.o0 = s.y->d, y.a0 = s.y->d, y.f0 = s.dy->d, y.n = s.y->len;
        v.push_back(x);
        v.push_back(y);
    }));
    // spmv_t_primal(y, x, τ → aty, x_t, x_bar); spmv_dual(x_bar, y, σ → ax, y_t)
    P.spmv(b->k_cspmv_primal, b->k_cfin_primal, each([](std::vector<COp>& v, const StepWork& s) {
               COp o;
               o.in = s.y->d, o.a0 = s.x->d, o.a1 = s.problem->c, o.a2 = s.problem->lv, o.a3 = s.problem->uv;
               o.s0 = s.tau, o.o0 = s.aty->d, o.o1 = s.x_t->d, o.o2 = s.x_bar->d;
               v.push_back(o);
           }),
           probs, false);
    P.spmv(b->k_cspmv_dual, b->k_cfin_dual, each([](std::vector<COp>& v, const StepWork& s) {
               COp o;
               o.in = s.x_bar->d, o.a0 = s.y->d, o.a1 = s.problem->lc, o.a2 = s.problem->uc, o.s0 = s.sigma;
               o.o0 = s.ax->d, o.o1 = s.y_t->d;
               v.push_back(o);
           }),
           probs, true);
    using S = Solver<DriverBackend>;
    if (!check) {
        // The first iteration of a restart epoch: r(z_anchor)'s sums
        // (enqueue_fixed_point), then the Halpern steps.
        P.elementwise(b->k_csub, each([](std::vector<COp>& v, const StepWork& s) {
            COp x, y;
            x.a0 = s.x_t->d, x.a1 = s.x->d, x.o0 = s.tmp_n->d, x.n = s.x->len;
            y.a0 = s.y_t->d, y.a1 = s.y->d, y.o0 = s.tmp_m2->d, y.n = s.y->len;
            v.push_back(x);
            v.push_back(y);
        }));
        P.spmv(b->k_cspmv_store, b->k_cfin_store, each([](std::vector<COp>& v, const StepWork& s) {
                   COp o;
                   o.in = s.tmp_n->d, o.o0 = s.tmp_m->d;
                   v.push_back(o);
               }),
               probs, true);
        P.reduce_any(each([&](std::vector<COp>& v, const StepWork& s) {
                         COp d0, d1, d2;
                         d0.a0 = s.tmp_n->d, d0.a1 = s.tmp_n->d, d0.n = s.tmp_n->len, d0.dst = slot(s, s.fix_base);
                         d1.a0 = s.tmp_m2->d, d1.a1 = s.tmp_m2->d, d1.n = s.tmp_m2->len;
                         d1.dst = slot(s, s.fix_base + 1);
                         d2.a0 = s.tmp_m2->d, d2.a1 = s.tmp_m->d, d2.n = s.tmp_m2->len;
                         d2.dst = slot(s, s.fix_base + 2);
                         for (COp* d : {&d0, &d1, &d2}) {
                             d->kind = kRedDot;
                             v.push_back(*d);
                         }
                     }),
                     h);
        enqueue_halpern(P, b, ws);
        return;
    }
    // A check iteration (Solver::enqueue_check): the four differences, the
    // five products, every reduction in one launch, the Halpern steps.
    P.elementwise(b->k_csub, each([](std::vector<COp>& v, const StepWork& s) {
        const std::pair<const sp_vec*, const sp_vec*> ab[4] = {
            {s.x_t, s.x}, {s.y_t, s.y}, {s.y_t, s.y_anchor}, {s.x_t, s.x_anchor}};
        sp_vec* out[4] = {s.tmp_n, s.tmp_m2, s.dya, s.dxa};
        for (int q = 0; q < 4; ++q) {
            COp o;
            o.a0 = ab[q].first->d, o.a1 = ab[q].second->d, o.o0 = out[q]->d, o.n = out[q]->len;
            v.push_back(o);
        }
    }));
    auto product = [&](auto in, auto out, bool a) {
        P.spmv(b->k_cspmv_store, b->k_cfin_store, each([&](std::vector<COp>& v, const StepWork& s) {
                   COp o;
                   o.in = in(s)->d, o.o0 = out(s)->d;
                   v.push_back(o);
               }),
               probs, a);
    };
    product([](const StepWork& s) { return s.tmp_n; }, [](const StepWork& s) { return s.tmp_m; }, true);
    product([](const StepWork& s) { return s.x_t; }, [](const StepWork& s) { return s.axt; }, true);
    product([](const StepWork& s) { return s.y_t; }, [](const StepWork& s) { return s.atyt; }, false);
    product([](const StepWork& s) { return s.dya; }, [](const StepWork& s) { return s.atdya; }, false);
    product([](const StepWork& s) { return s.dxa; }, [](const StepWork& s) { return s.adxa; }, true);
    P.reduce_any(each([&](std::vector<COp>& v, const StepWork& s) {
                     const sp_problem* p = s.problem;
                     auto red = [&](RedKind kind, std::vector<const double*> a, bool rows_and_cols, int64_t len,
                                    int at) {
                         COp o;
                         const double** f[] = {&o.a0, &o.a1, &o.a2, &o.a3, &o.a4, &o.a5, &o.a6, &o.a7};
                         for (size_t i = 0; i < a.size(); ++i) *f[i] = a[i];
                         o.kind = kind;
                         o.m = rows_and_cols ? p->m : 0;
                         o.n = rows_and_cols ? p->n : len;
                         o.dst = slot(s, at);
                         v.push_back(o);
                     };
                     const int64_t m = p->m, n = p->n;
                     red(kRedDot, {s.tmp_n->d, s.tmp_n->d}, false, n, s.fix_base);
                     red(kRedDot, {s.tmp_m2->d, s.tmp_m2->d}, false, m, s.fix_base + 1);
                     red(kRedDot, {s.tmp_m2->d, s.tmp_m->d}, false, m, s.fix_base + 2);
                     red(kRedPrimal, {s.axt->d, p->lc, p->uc, p->pf, p->c, s.x_t->d}, true, 0, S::slot_primal());
                     red(kRedDual, {s.y_t->d, p->lc, p->uc, p->c, s.atyt->d, p->lv, p->uv, p->df}, true, 0,
                         S::slot_dual());
                     red(kRedStep, {s.x->d, s.x_t->d, s.aty->d, s.atyt->d, p->df}, false, n, S::slot_step());
                     v.back().s0 = 1.0 / s.tau;
                     if (s.step_implied) {
                         red(kRedDual, {s.y_t->d, p->lc, p->uc, p->c, s.c->d, p->lv, p->uv, p->df}, true, 0,
                             S::slot_step_rows());
                         const double* pairs[4][2] = {{s.c->d, s.x_t->d}, {s.aty->d, s.x_t->d},
                                                      {s.x_t->d, s.x_t->d}, {s.x->d, s.x_t->d}};
                         for (int q = 0; q < 4; ++q)
                             red(kRedDot, {pairs[q][0], pairs[q][1]}, false, n, S::slot_step_dots() + q);
                     }
                     red(kRedDualRay, {s.dya->d, p->lc, p->uc, s.atdya->d, p->lv, p->uv}, true, 0,
                         S::slot_dual_ray());
                     red(kRedPrimalRay, {s.adxa->d, p->lc, p->uc, s.dxa->d, p->c, p->lv, p->uv}, true, 0,
                         S::slot_primal_ray());
                     red(kRedDist, {s.x_t->d, s.x_anchor->d}, false, n, S::slot_dist_x());
                     red(kRedDist, {s.y_t->d, s.y_anchor->d}, false, m, S::slot_dist_y());
                 }),
                 h);
    enqueue_halpern(P, b, ws);
}

// Every entry's result slots into one buffer (gathered[i·count + j]), for
// one readback.
inline void gather_slots(Phase& P, const Hip& h, const std::vector<StepWork>& ws, double* gathered, int count) {
    std::vector<COp> v;
    for (size_t i = 0; i < ws.size(); ++i) {
        COp o;
        o.a0 = ws[i].slots, o.o0 = gathered + i * (size_t)count, o.n = count;
        v.push_back(o);
    }
    P.elementwise(h.b->k_ccopy, v);
}

// Solver::enter_plain for every entry: the base's FP64 products, zero
// displacements, and the anchors' offsets from the base.
inline void enqueue_plain_entries(Phase& P, const Hip& h, const std::vector<PlainEntry>& es) {
    if (es.empty()) return;
    sp_backend* b = h.b;
    std::vector<COp> t, a;
    std::vector<sp_problem*> pt;
    for (const PlainEntry& e : es) {
        COp o;
        o.in = e.y->d, o.o0 = e.atya->d;
        t.push_back(o);
        pt.push_back(e.problem);
    }
    // spmv_t(y → atya), then spmv(x → axa), as enter_plain.
    P.spmv(b->k_cspmv_store, b->k_cfin_store, t, pt, false);
    for (const PlainEntry& e : es) {
        COp o;
        o.in = e.x->d, o.o0 = e.axa->d;
        a.push_back(o);
    }
    P.spmv(b->k_cspmv_store, b->k_cfin_store, a, pt, true);
    std::vector<COp> d;
    for (const PlainEntry& e : es) {
        COp x, y, ex, ey;
        x.a0 = e.x->d, x.a1 = e.x->d, x.f0 = e.dx->d, x.n = e.x->len;
        y.a0 = e.y->d, y.a1 = e.y->d, y.f0 = e.dy->d, y.n = e.y->len;
        ex.a0 = e.x_anchor->d, ex.a1 = e.x->d, ex.f0 = e.ex->d, ex.n = e.x->len;
        ey.a0 = e.y_anchor->d, ey.a1 = e.y->d, ey.f0 = e.ey->d, ey.n = e.y->len;
        d.push_back(x);
        d.push_back(y);
        d.push_back(ex);
        d.push_back(ey);
    }
    P.elementwise(b->k_cdelta_from, d);
}

using PowerWork = Solver<DriverBackend>::PowerWork;

// `rounds` power-iteration rounds of every entry (Solver::start's rounds):
// tmp_m = A'v, v = A'ᵀtmp_m, slot = v·v, v *= 1/sqrt(slot).
inline void enqueue_power_rounds(Phase& P, const Hip& h, const std::vector<PowerWork>& ws, int rounds) {
    if (ws.empty()) return;
    sp_backend* b = h.b;
    std::vector<sp_problem*> probs;
    for (const PowerWork& w : ws) probs.push_back(w.problem);
    constexpr int base = Solver<DriverBackend>::slot_power();
    for (int r = 0; r < rounds; ++r) {
        std::vector<COp> a, t, d, s;
        for (const PowerWork& w : ws) {
            COp o;
            o.in = w.v->d, o.o0 = w.tmp_m->d;
            a.push_back(o);
            COp q;
            q.in = w.tmp_m->d, q.o0 = w.v->d;
            t.push_back(q);
            COp dd;
            dd.kind = kRedDot;
            dd.a0 = w.v->d, dd.a1 = w.v->d, dd.n = w.v->len, dd.dst = w.slots + base + r;
            d.push_back(dd);
            COp sc;
            sc.o0 = w.v->d, sc.a0 = w.slots + base + r, sc.n = w.v->len;
            s.push_back(sc);
        }
        P.spmv(b->k_cspmv_store, b->k_cfin_store, a, probs, true);
        P.spmv(b->k_cspmv_store, b->k_cfin_store, t, probs, false);
        P.reduce_any(d, h);
        P.elementwise(b->k_cscale_inv_sqrt, s);
    }
}

// The restart copies of a check phase, in one launch.
inline void enqueue_copies(Phase& P, const Hip& h, const std::vector<std::pair<sp_vec*, sp_vec*>>& copies) {
    std::vector<COp> v;
    for (const auto& c : copies) {
        COp o;
        o.a0 = c.first->d, o.o0 = c.second->d, o.n = c.first->len;
        v.push_back(o);
    }
    P.elementwise(h.b->k_ccopy, v);
}

}  // namespace sp
