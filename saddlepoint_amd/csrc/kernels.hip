// Hand-written gfx950 kernels for the saddlepoint PDHG backend.
//
// Built offline into a code object (hipcc --genco --offload-arch=gfx950) and
// loaded with hipModuleLoad by the host library; every kernel is extern "C".
// Each one is one pass over HBM. All of the work is bandwidth-bound, so there
// is no MFMA here: the design goals are coalesced streaming of the CSR arrays,
// one gather of the input vector per nonzero, and epilogues fused into the
// SpMV so that vectors are read and written once per iteration.
//
// Numerics follow the reference CPU backend (crates/saddlepoint-cpu/src/lib.rs):
// the build uses -ffp-contract=off (no fused multiply-add, as in Rust), clamp
// has Rust's f64::clamp semantics, and every reduction has a fixed order that
// depends only on the launch shape, so repeated calls give identical bits.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef long long i64;

#define WG 256
#define WAVE 64
#define NWAVES (WG / WAVE)

// Stream blocks stage at most this many products in LDS. A longer row gets
// workgroups of its own, one per SP_CHUNK nonzeros.
#define SP_NNZ_F64 2048
#define SP_NNZ_F32 4096
#define SP_CHUNK 4096   // nonzeros of a long row per workgroup
#define SP_BATCH 4096   // nonzeros a workgroup loads at once
// Hot tables (spmv_body_hot): entries of the gathered vector kept in LDS.
#define SP_HOT_F64 2048
#define SP_HOT_F32 4096

// One SpMV work item. Stream block: rows [r0, r1) whose nonzeros are [k0, k1).
// Long-row chunk: row r0, nonzeros [k0, k1) of it, r1 = -1 - partial slot.
struct Desc {
    int r0, r1, k0, k1;
};

// ---- scalar helpers (cpu/src/lib.rs:70-99, 262-269) -------------------------

template <class T>
__device__ __forceinline__ T clampv(T v, T lo, T hi) {
    // Rust f64::clamp: NaN passes through (not fmin/fmax, which drop it).
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
// Epilogue stores. Plain stores, like the plain loads of the matrix streams
// below: on gfx950 the non-temporal hint on either made the config-1 SpMVs
// slower (A side 51 -> 60 us, A'ᵀ side 55 -> 58 us) and config 2 no faster.
template <class T>
__device__ __forceinline__ void st(T* p, T v) {
    *p = v;
}

__device__ __forceinline__ double excess(double v, double lo, double hi) {
    return v < lo ? v - lo : (v > hi ? v - hi : 0.0);
}
__device__ __forceinline__ double bound_term(double v, double lo, double hi) {
    if (v > 0.0 && __builtin_isfinite(lo)) return v * lo;
    if (v < 0.0 && __builtin_isfinite(hi)) return v * hi;
    return 0.0;
}
__device__ __forceinline__ double forbidden(double v, double lo, double hi) {
    return ((v > 0.0 && lo == -__builtin_inf()) || (v < 0.0 && hi == __builtin_inf())) ? v : 0.0;
}
__device__ __forceinline__ double recession_violation(double v, double lo, double hi) {
    const double l = __builtin_isfinite(lo) ? 0.0 : -__builtin_inf();
    const double h = __builtin_isfinite(hi) ? 0.0 : __builtin_inf();
    return excess(v, l, h);
}

// ---- fixed-order block reductions -------------------------------------------

template <class A>
__device__ __forceinline__ A wave_sum(A s) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, WAVE);
    return s;
}

// Sums K values over the 256 threads; the result is valid in thread 0. The
// order (a butterfly inside each wave, then waves 0..3 in turn) is fixed.
template <int K, class A>
__device__ __forceinline__ void block_sum(A (&s)[K]) {
    __shared__ A sh[K][NWAVES];
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        A v = wave_sum(s[k]);
        if (lane == 0) sh[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            A v = sh[k][0];
#pragma unroll
            for (int w = 1; w < NWAVES; ++w) v += sh[k][w];
            s[k] = v;
        }
    }
    __syncthreads();
}

// ---- SpMV with a fused per-row epilogue --------------------------------------
//
// out_row = Σ_k val[k] * x[col[k]] over the row, then the epilogue on the sum.
//
// Stream blocks (rows [r0, r1) holding at most NNZ nonzeros and at most
// SP_ROWS rows) stage the products in LDS with fully coalesced loads, then
// reduce rows with a merge-path split (Merrill & Garland, "Merge-based
// parallel sparse matrix-vector multiplication", SC'16): the merged sequence
// of row ends and nonzeros is cut into WG equal pieces, so every thread does
// the same work whatever the row lengths. A row that crosses pieces gets the
// partial sums of the preceding pieces added in order by the thread that
// finishes it, so the summation order is fixed for a given matrix and plan.
// Every load that does not depend on another one is issued up front: column
// indices, values, row pointers and the epilogue's per-row operands, then the
// gathers.
//
// Long rows are summed by a whole workgroup per chunk of SP_CHUNK nonzeros; a
// row of several chunks publishes the chunk sums and the finishing kernel
// adds them in chunk order.
#define SP_RPT 4                // rows per thread in the epilogue
#define SP_ROWS (SP_RPT * WG)   // most rows in one stream block

// Where the gathers of a batch read x[c[u]]: HBM (through the caches), or,
// for the H entries of a hot table, an LDS copy of x[0, H). The value is the
// same either way, so the choice never changes a result's bits.
struct GatherGlobal {
    template <int N, class X>
    __device__ __forceinline__ void batch(const X* __restrict__ x, const int (&c)[N], X (&v)[N]) const {
#pragma unroll
        for (int u = 0; u < N; ++u) v[u] = x[c[u]];
    }
};

template <class X, int H>
struct GatherHot {
    static_assert((H & (H - 1)) == 0, "H is a power of two");
    const X* hot;  // LDS: x[0, H)
    template <int N>
    __device__ __forceinline__ void batch(const X* __restrict__ x, const int (&c)[N], X (&v)[N]) const {
        // Every lane reads LDS (cheap), then only the lanes outside the hot
        // set issue a memory request: each load runs with them alone in the
        // wave's mask, and a wave with none skips it. The empty asm makes the
        // LDS values opaque (one wait for all of them), so the compiler
        // cannot merge the two reads of an element into one load through a
        // generic (flat) pointer, which would wait on both memory counters.
#pragma unroll
        for (int u = 0; u < N; ++u) v[u] = hot[c[u] & (H - 1)];
#pragma unroll
        for (int u = 0; u < N; ++u) asm volatile("" : "+v"(v[u]));
#pragma unroll
        for (int u = 0; u < N; ++u)
            if (c[u] >= H) v[u] = x[c[u]];
    }
};

template <int NNZ, class V, class X, class A, class Epi, class G = GatherGlobal>
__device__ __forceinline__ void spmv_block(const Desc d, const int* __restrict__ rp, const int* __restrict__ ci,
                                           const V* __restrict__ va, const X* __restrict__ x,
                                           A* __restrict__ partials, const Epi& epi, const G& gat = G{}) {
    __shared__ A prod[NNZ];
    __shared__ A sums[SP_ROWS];
    __shared__ int rend[SP_ROWS];
    __shared__ int crow[WG];
    __shared__ A cval[WG];
    const int tid = threadIdx.x;
    if (d.r1 >= 0) {
        const int r0 = d.r0, nr = d.r1 - d.r0, k0 = d.k0, n = d.k1 - d.k0;
        constexpr int PER = NNZ / WG;
        // Straight-line loads: out-of-range slots read a clamped, in-bounds
        // address and are masked at use, so no load sits behind a branch and
        // the compiler keeps them all in flight.
        const int last = n > 0 ? n - 1 : 0;
        const int kb = n > 0 ? k0 : 0;
        int c[PER];
        V v[PER];
        if (n > 0) {
#pragma unroll
            for (int u = 0; u < PER; ++u) c[u] = ci[kb + min(tid + u * WG, last)];
#pragma unroll
            for (int u = 0; u < PER; ++u) v[u] = va[kb + min(tid + u * WG, last)];
        }
        int re[SP_RPT];
        typename Epi::Pre pre[SP_RPT];
#pragma unroll
        for (int j = 0; j < SP_RPT; ++j) {
            if (j * WG < nr) {  // uniform across the workgroup
                const int rr = r0 + min(tid + j * WG, nr - 1);
                re[j] = rp[rr + 1];
                pre[j] = epi.load(rr);
            }
        }
        if (n > 0) {
            X xv[PER];
            gat.batch(x, c, xv);
#pragma unroll
            for (int u = 0; u < PER; ++u)
                if (tid + u * WG < n) prod[tid + u * WG] = (A)v[u] * (A)xv[u];
        }
#pragma unroll
        for (int j = 0; j < SP_RPT; ++j)
            if (j * WG < nr && tid + j * WG < nr) rend[tid + j * WG] = re[j] - k0;
        __syncthreads();

        // Merge path: this thread's piece of the nr + n merged items.
        const int L = nr + n;
        const int d0 = (int)(((long long)tid * L) / WG), d1 = (int)(((long long)(tid + 1) * L) / WG);
        int lo = max(0, d0 - n), hi = min(d0, nr);
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (rend[mid] <= d0 - mid - 1) lo = mid + 1;
            else hi = mid;
        }
        int i = lo, j = d0 - lo;
        int first = -1;  // the first row this thread finishes (it may have begun earlier)
        A acc = 0;
        int e = i < nr ? rend[i] : 0;
        for (int step = d0; step < d1; ++step) {
            if (i < nr && j < e) {
                acc += prod[j];
                ++j;
            } else if (i < nr) {
                if (first < 0) first = i;
                sums[i] = acc;
                acc = 0;
                ++i;
                e = i < nr ? rend[i] : 0;
            }
        }
        crow[tid] = i < nr ? i : -1;  // the row left open, with its partial sum
        cval[tid] = acc;
        __syncthreads();
        if (first >= 0 && tid > 0 && crow[tid - 1] == first) {
            // The partial sums of row `first` left open by the threads before
            // this one, in thread order. They start at the thread whose piece
            // holds the row's first nonzero, merged position P = first + (its
            // offset): the last thread t with t·L/WG <= P. (Threads before it
            // that end with this row open have added nothing to it: their +0
            // leaves the sum's bits unchanged, so they need not be walked.)
            const long long P = (long long)first + (first > 0 ? rend[first - 1] : 0);
            int q = (int)(((P + 1) * WG - 1) / L);
            A t = 0;
            for (; q < tid; ++q) t += cval[q];
            sums[first] = t + sums[first];
        }
        __syncthreads();
#pragma unroll
        for (int jj = 0; jj < SP_RPT; ++jj) {
            if (jj * WG >= nr) break;  // uniform
            const int rr = tid + jj * WG;
            if (rr < nr) epi.apply(r0 + rr, sums[rr], pre[jj]);
        }
    } else {
        const int r = d.r0, pslot = -1 - d.r1, k0 = d.k0, k1 = d.k1;
        // Up to SP_CHUNK nonzeros of one row, in batches of SP_BATCH with the
        // same straight-line pattern as the stream blocks; each thread keeps
        // its running sum in a register across batches.
        constexpr int PERL = SP_BATCH / WG;
        A s[1] = {0};
        for (int base = k0; base < k1; base += SP_BATCH) {
            const int n = min(SP_BATCH, k1 - base), lastl = n - 1;
            int c[PERL];
            V v[PERL];
#pragma unroll
            for (int u = 0; u < PERL; ++u) c[u] = ci[base + min(tid + u * WG, lastl)];
#pragma unroll
            for (int u = 0; u < PERL; ++u) v[u] = va[base + min(tid + u * WG, lastl)];
            X xv[PERL];
            gat.batch(x, c, xv);
#pragma unroll
            for (int u = 0; u < PERL; ++u)
                if (tid + u * WG < n) s[0] += (A)v[u] * (A)xv[u];
        }
        block_sum<1>(s);
        const int rs = rp[r], len = rp[r + 1] - rs;
        const int nch = (len + SP_CHUNK - 1) / SP_CHUNK;
        if (nch == 1) {
            if (tid == 0) epi.apply(r, s[0], epi.load(r));
            return;
        }
        // A row of several chunks: publish this chunk's sum; a finishing
        // kernel adds the chunks in order and applies the epilogue.
        if (tid == 0) partials[pslot] = s[0];
    }
}

template <int NNZ, class V, class X, class A, class Epi>
__device__ __forceinline__ void spmv_body(const Desc* __restrict__ desc, const int* __restrict__ rp,
                                          const int* __restrict__ ci, const V* __restrict__ va,
                                          const X* __restrict__ x, A* __restrict__ partials, const Epi& epi) {
    spmv_block<NNZ, V, X, A, Epi>(desc[blockIdx.x], rp, ci, va, x, partials, epi);
}

// The same work items with a hot table: each workgroup copies x[0, H) into
// LDS once and then takes work items ndesc apart (a grid of a few workgroups
// per CU, so the copy is paid a few hundred times, not once per item). For a
// matrix whose column indices are mostly below H (A'ᵀ once the rows of A' are
// relabeled long rows first: the first rows of A' are the most referenced
// entries of y), most gathers then read LDS and only the rest go to memory.
// Each item is computed exactly as by spmv_body, so the sums are the same.
template <int NNZ, int H, class V, class X, class A, class Epi>
__device__ __forceinline__ void spmv_body_hot(const Desc* __restrict__ desc, int ndesc, const int* __restrict__ rp,
                                              const int* __restrict__ ci, const V* __restrict__ va,
                                              const X* __restrict__ x, A* __restrict__ partials, const Epi& epi) {
    __shared__ X hot[H];
    for (int i = threadIdx.x; i < H; i += WG) hot[i] = x[i];  // the host launches this only when len(x) >= H
    __syncthreads();
    const GatherHot<X, H> g{hot};
    for (int b = blockIdx.x; b < ndesc; b += gridDim.x) {
        spmv_block<NNZ, V, X, A, Epi>(desc[b], rp, ci, va, x, partials, epi, g);
        __syncthreads();  // the next item reuses the block's LDS
    }
}

// The finishing pass for rows of several chunks: one wave per row adds the
// row's chunk sums in a fixed order and applies the epilogue. (Combining in
// the SpMV itself would need an agent-scope release fence per workgroup, and
// on this part, with one L2 per XCD, that is an L2 write-back each time.)
// A row whose sum is split into nch partial sums: they are
// partials[pbase + q·stride] for q < nch (stride 0 reads as 1).
struct LongRow {
    int row, pbase, nch, stride;
};

template <class A, class Epi>
__device__ __forceinline__ void finish_body(const LongRow* __restrict__ longs, int nlong,
                                            const A* __restrict__ partials, const Epi& epi, int w) {
    const int lane = threadIdx.x % WAVE;
    if (w >= nlong) return;  // uniform per wave
    const LongRow L = longs[w];
    const i64 stride = L.stride ? L.stride : 1;
    // The row's epilogue operands are loaded beside its partial sums (every
    // lane, one address), not after the sum: one memory round trip less.
    const typename Epi::Pre pre = epi.load(L.row);
    A s = 0;
    for (int q = lane; q < L.nch; q += WAVE) s += partials[L.pbase + q * stride];
    s = wave_sum(s);
    if (lane == 0) epi.apply(L.row, s, pre);
}

template <class A, class Epi>
__device__ __forceinline__ void finish_body(const LongRow* __restrict__ longs, int nlong,
                                            const A* __restrict__ partials, const Epi& epi) {
    finish_body<A>(longs, nlong, partials, epi, (int)((blockIdx.x * WG + threadIdx.x) / WAVE));
}

// Epilogues: load(row) fetches the row's operands (issued early), apply(row,
// sum, operands) computes and stores.
template <class A>
struct EpiStore {
    A* out;
    struct Pre {};
    __device__ Pre load(int) const { return {}; }
    __device__ void apply(int r, A s, const Pre&) const { st(out + r, s); }
};

// spmv_t + primal_step: aty = A'ᵀy; x_t = clamp(x - τ(c - aty), lv, uv);
// x_bar = 2 x_t - x (cpu/src/lib.rs:147-160). Stores aty, x_t and x_bar.
struct EpiPrimal {
    const double *x, *c, *lv, *uv;
    double tau;
    double *aty, *x_t, *x_bar;
    struct Pre {
        double x, c, lv, uv;
    };
    __device__ Pre load(int j) const { return {x[j], c[j], lv[j], uv[j]}; }
    __device__ void apply(int j, double s, const Pre& q) const {
        const double t = clampv(q.x - tau * (q.c - s), q.lv, q.uv);
        st(aty + j, s);
        st(x_t + j, t);
        st(x_bar + j, 2.0 * t - q.x);
    }
};

// As EpiPrimal, then the reflected Halpern update of x in place
// (cpu/src/lib.rs:177-187); x_t and aty are not stored.
struct EpiPrimalHalpern {
    double* x;
    const double *c, *lv, *uv, *anchor;
    double tau, w, rho;
    double* x_bar;
    struct Pre {
        double x, c, lv, uv, anchor;
    };
    __device__ Pre load(int j) const { return {x[j], c[j], lv[j], uv[j], anchor[j]}; }
    __device__ void apply(int j, double s, const Pre& q) const {
        const double t = clampv(q.x - tau * (q.c - s), q.lv, q.uv);
        st(x_bar + j, 2.0 * t - q.x);
        st(x + j, w * ((1.0 + rho) * t - rho * q.x) + (1.0 - w) * q.anchor);
    }
};

// spmv + dual_step: ax = A'x_bar; y_t = y - σ(ax - clamp(ax - y/σ, lc, uc))
// (cpu/src/lib.rs:162-175). Stores ax and y_t.
struct EpiDual {
    const double *y, *lc, *uc;
    double sigma;
    double *ax, *y_t;
    struct Pre {
        double y, lc, uc;
    };
    __device__ Pre load(int i) const { return {y[i], lc[i], uc[i]}; }
    __device__ void apply(int i, double s, const Pre& q) const {
        const double v = s - q.y / sigma;
        st(ax + i, s);
        st(y_t + i, q.y - sigma * (s - clampv(v, q.lc, q.uc)));
    }
};

struct EpiDualHalpern {
    double* y;
    const double *lc, *uc, *anchor;
    double sigma, w, rho;
    struct Pre {
        double y, lc, uc, anchor;
    };
    __device__ Pre load(int i) const { return {y[i], lc[i], uc[i], anchor[i]}; }
    __device__ void apply(int i, double s, const Pre& q) const {
        const double v = s - q.y / sigma;
        const double t = q.y - sigma * (s - clampv(v, q.lc, q.uc));
        st(y + i, w * ((1.0 + rho) * t - rho * q.y) + (1.0 - w) * q.anchor);
    }
};

#define SPMV_ARGS                                                                                    \
    const Desc *__restrict__ desc, const int *__restrict__ rp, const int *__restrict__ ci,          \
        const double *__restrict__ va, const double *__restrict__ xin, double *__restrict__ partials

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_f64(SPMV_ARGS, double* __restrict__ out) {
    spmv_body<SP_NNZ_F64, double, double, double>(desc, rp, ci, va, xin, partials, EpiStore<double>{out});
}

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_t_primal_f64(SPMV_ARGS, const double* __restrict__ x,
                                                                      const double* __restrict__ c,
                                                                      const double* __restrict__ lv,
                                                                      const double* __restrict__ uv, double tau,
                                                                      double* __restrict__ aty,
                                                                      double* __restrict__ x_t,
                                                                      double* __restrict__ x_bar) {
    spmv_body<SP_NNZ_F64, double, double, double>(desc, rp, ci, va, xin, partials,
                                                  EpiPrimal{x, c, lv, uv, tau, aty, x_t, x_bar});
}

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_t_primal_halpern_f64(
    SPMV_ARGS, double* __restrict__ x, const double* __restrict__ c, const double* __restrict__ lv,
    const double* __restrict__ uv, const double* __restrict__ anchor, double tau, double w, double rho,
    double* __restrict__ x_bar) {
    spmv_body<SP_NNZ_F64, double, double, double>(desc, rp, ci, va, xin, partials,
                                                  EpiPrimalHalpern{x, c, lv, uv, anchor, tau, w, rho, x_bar});
}

// A'ᵀ side of a plain iteration with a hot table of y (spmv_body_hot): 2,048
// entries (16 KB, 47 KB of LDS in all, three workgroups per CU). Measured on
// a 4M x 10M matrix of Zipf(1.5) row lengths, rows long rows first: a larger
// table (two workgroups per CU) gains nothing, a smaller one at four per CU
// slightly less.
extern "C" __global__ __launch_bounds__(WG) void sp_spmv_t_primal_halpern_hot_f64(
    SPMV_ARGS, int ndesc, double* __restrict__ x, const double* __restrict__ c, const double* __restrict__ lv,
    const double* __restrict__ uv, const double* __restrict__ anchor, double tau, double w, double rho,
    double* __restrict__ x_bar) {
    spmv_body_hot<SP_NNZ_F64, SP_HOT_F64, double, double, double>(
        desc, ndesc, rp, ci, va, xin, partials, EpiPrimalHalpern{x, c, lv, uv, anchor, tau, w, rho, x_bar});
}

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_dual_f64(SPMV_ARGS, const double* __restrict__ y,
                                                                  const double* __restrict__ lc,
                                                                  const double* __restrict__ uc, double sigma,
                                                                  double* __restrict__ ax,
                                                                  double* __restrict__ y_t) {
    spmv_body<SP_NNZ_F64, double, double, double>(desc, rp, ci, va, xin, partials,
                                                  EpiDual{y, lc, uc, sigma, ax, y_t});
}

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_dual_halpern_f64(
    SPMV_ARGS, double* __restrict__ y, const double* __restrict__ lc, const double* __restrict__ uc,
    const double* __restrict__ anchor, double sigma, double w, double rho) {
    spmv_body<SP_NNZ_F64, double, double, double>(desc, rp, ci, va, xin, partials,
                                                  EpiDualHalpern{y, lc, uc, anchor, sigma, w, rho});
}


#define FIN_ARGS(A) const LongRow *__restrict__ longs, int nlong, const A *__restrict__ partials

extern "C" __global__ __launch_bounds__(WG) void sp_finish_f64(FIN_ARGS(double), double* __restrict__ out) {
    finish_body<double>(longs, nlong, partials, EpiStore<double>{out});
}

extern "C" __global__ __launch_bounds__(WG) void sp_finish_t_primal_f64(FIN_ARGS(double), const double* __restrict__ x,
                                                                        const double* __restrict__ c,
                                                                        const double* __restrict__ lv,
                                                                        const double* __restrict__ uv, double tau,
                                                                        double* __restrict__ aty,
                                                                        double* __restrict__ x_t,
                                                                        double* __restrict__ x_bar) {
    finish_body<double>(longs, nlong, partials, EpiPrimal{x, c, lv, uv, tau, aty, x_t, x_bar});
}

extern "C" __global__ __launch_bounds__(WG) void sp_finish_t_primal_halpern_f64(
    FIN_ARGS(double), double* __restrict__ x, const double* __restrict__ c, const double* __restrict__ lv,
    const double* __restrict__ uv, const double* __restrict__ anchor, double tau, double w, double rho,
    double* __restrict__ x_bar) {
    finish_body<double>(longs, nlong, partials, EpiPrimalHalpern{x, c, lv, uv, anchor, tau, w, rho, x_bar});
}

extern "C" __global__ __launch_bounds__(WG) void sp_finish_dual_f64(FIN_ARGS(double), const double* __restrict__ y,
                                                                    const double* __restrict__ lc,
                                                                    const double* __restrict__ uc, double sigma,
                                                                    double* __restrict__ ax, double* __restrict__ y_t) {
    finish_body<double>(longs, nlong, partials, EpiDual{y, lc, uc, sigma, ax, y_t});
}

extern "C" __global__ __launch_bounds__(WG) void sp_finish_dual_halpern_f64(
    FIN_ARGS(double), double* __restrict__ y, const double* __restrict__ lc, const double* __restrict__ uc,
    const double* __restrict__ anchor, double sigma, double w, double rho) {
    finish_body<double>(longs, nlong, partials, EpiDualHalpern{y, lc, uc, anchor, sigma, w, rho});
}

// ---- column panels: the long rows of a wide matrix ------------------------------
//
// A row's gathers land anywhere in x, and when x is far larger than an XCD's
// L2 (config 2: 80 MB against 4 MB) almost every gather misses it. The long
// rows of such a matrix (more than 64 nonzeros; on config 2 they hold 94% of
// the nonzeros) are therefore also kept as a panel matrix: its row
// q = p·L + l holds long row l's nonzeros whose columns fall in panel p
// (columns [p·W, (p+1)·W), a contiguous part of the row since its columns are
// sorted), and its rows are stored and dispatched panel by panel. The
// workgroups in flight then all gather from one panel of x, a slice that fits
// in L2. The panel matrix is an ordinary CSR matrix with its own SpMV plan;
// its row sums are stored in row order, so long row l's P panel sums sit at
// l, l + L, ..., and the long-row finishing kernels (stride L) add them in
// panel order and apply the epilogue. The short rows run as stream blocks of the original
// matrix. Every A' product uses this one decomposition, so the summation
// order is a function of the matrix and its plan, as before.

// First position k in [a, b) with ci[k] >= v (ci sorted on [a, b)).
__device__ __forceinline__ int lower_bound_col(const int* __restrict__ ci, int a, int b, i64 v) {
    while (a < b) {
        const int mid = a + ((b - a) >> 1);
        if ((i64)ci[mid] < v) a = mid + 1;
        else b = mid;
    }
    return a;
}

// For each panel-matrix row q = p·L + l: where long row rows[l]'s panel-p
// nonzeros start in A and how many there are.
extern "C" __global__ __launch_bounds__(WG) void sp_panel_count(const int* __restrict__ rp, const int* __restrict__ ci,
                                                                const int* __restrict__ rows, int L, int P, int shift,
                                                                int* __restrict__ start, int* __restrict__ count) {
    const i64 q = (i64)blockIdx.x * WG + threadIdx.x;
    if (q >= (i64)P * L) return;
    const int p = (int)(q / L), r = rows[q % L];
    const int lo = lower_bound_col(ci, rp[r], rp[r + 1], (i64)p << shift);
    const int hi = lower_bound_col(ci, lo, rp[r + 1], (i64)(p + 1) << shift);
    start[q] = lo;
    count[q] = hi - lo;
}

// The panel matrix's column indices and values (one wave per row).
extern "C" __global__ __launch_bounds__(WG) void sp_panel_gather(const int* __restrict__ start,
                                                                 const int* __restrict__ brp, i64 nq,
                                                                 const int* __restrict__ ci,
                                                                 const double* __restrict__ va,
                                                                 int* __restrict__ bci, double* __restrict__ bva) {
    const i64 q = ((i64)blockIdx.x * WG + threadIdx.x) / WAVE;
    const int lane = threadIdx.x % WAVE;
    if (q >= nq) return;  // uniform per wave
    const int s = start[q], d0 = brp[q], len = brp[q + 1] - d0;
    for (int t = lane; t < len; t += WAVE) {
        bci[d0 + t] = ci[s + t];
        bva[d0 + t] = va[s + t];
    }
}

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_panel_f64(SPMV_ARGS, double* __restrict__ out) {
    spmv_body<SP_NNZ_F64, double, double, double>(desc, rp, ci, va, xin, partials, EpiStore<double>{out});
}

extern "C" __global__ __launch_bounds__(WG) void sp_finish_panel_f64(FIN_ARGS(double), double* __restrict__ out) {
    finish_body<double>(longs, nlong, partials, EpiStore<double>{out});
}

// Row relabeling: new row i is old row order[i], its nonzeros in their order
// (one wave per row).
extern "C" __global__ __launch_bounds__(WG) void sp_rows_gather(const int* __restrict__ rp_old,
                                                                const int* __restrict__ ci_old,
                                                                const double* __restrict__ va_old,
                                                                const int* __restrict__ rp_new,
                                                                const int* __restrict__ order, int m,
                                                                int* __restrict__ ci_new,
                                                                double* __restrict__ va_new) {
    const int i = (int)(((i64)blockIdx.x * WG + threadIdx.x) / WAVE), lane = threadIdx.x % WAVE;
    if (i >= m) return;  // uniform per wave
    const int src = order[i], s0 = rp_old[src], d0 = rp_new[i], len = rp_new[i + 1] - d0;
    for (int t = lane; t < len; t += WAVE) {
        ci_new[d0 + t] = ci_old[s0 + t];
        va_new[d0 + t] = va_old[s0 + t];
    }
}

// out[i] = in[idx[i]].
extern "C" __global__ __launch_bounds__(WG) void sp_gather_f64(const double* __restrict__ in,
                                                               const int* __restrict__ idx,
                                                               double* __restrict__ out, i64 n) {
    for (i64 i = (i64)blockIdx.x * WG + threadIdx.x; i < n; i += (i64)gridDim.x * WG) out[i] = in[idx[i]];
}

// ---- mixed precision: FP32 matrix and displacements, FP64 anchors ------------
//
// Between checks the iterates are held as z = z_base + dz: z_base is the FP64
// iterate at the last check and dz an FP32 displacement from it. With
// A'ᵀy = A'ᵀy_base + A'ᵀdy and A'x_bar = A'x_base + A'dx_bar, the FP64
// products of the base are formed when plain iterations resume after a check,
// and the per-iteration SpMVs run on FP32 values with FP32 sums. Steps,
// projections and the Halpern update are evaluated in FP64 on z_base + dz, so
// the rounding is relative to at most a check interval's motion, not to the
// iterate. In displacement form the Halpern update
// z ← w((1+ρ)z_t - ρz) + (1-w)z_anchor reads
// dz ← w((1+ρ)(z_t - z_base) - ρ dz) + (1-w)ez, with ez = z_anchor - z_base
// held in FP32 (zero when the base is the anchor, right after a restart).

struct EpiPrimalMixed {
    float* dx;
    const double *xa, *c, *lv, *uv, *atya;  // xa: the base x
    const float* ex;                         // x_anchor - xa
    double tau, w, rho;
    float* dx_bar;
    struct Pre {
        float dx, ex;
        double xa, c, lv, uv, atya;
    };
    __device__ Pre load(int j) const { return {dx[j], ex[j], xa[j], c[j], lv[j], uv[j], atya[j]}; }
    __device__ void apply(int j, float s, const Pre& q) const {
        const double d = (double)q.dx;
        const double aty = q.atya + (double)s;
        const double t = clampv(q.xa + d - tau * (q.c - aty), q.lv, q.uv) - q.xa;
        st(dx_bar + j, (float)(2.0 * t - d));
        st(dx + j, (float)(w * ((1.0 + rho) * t - rho * d) + (1.0 - w) * (double)q.ex));
    }
};

struct EpiDualMixed {
    float* dy;
    const double *ya, *lc, *uc, *axa;  // ya: the base y
    const float* ey;                    // y_anchor - ya
    double sigma, w, rho;
    struct Pre {
        float dy, ey;
        double ya, lc, uc, axa;
    };
    __device__ Pre load(int i) const { return {dy[i], ey[i], ya[i], lc[i], uc[i], axa[i]}; }
    __device__ void apply(int i, float s, const Pre& q) const {
        const double d = (double)q.dy;
        const double y = q.ya + d;
        const double ax = q.axa + (double)s;
        const double v = ax - y / sigma;
        const double t = y - sigma * (ax - clampv(v, q.lc, q.uc)) - q.ya;
        st(dy + i, (float)(w * ((1.0 + rho) * t - rho * d) + (1.0 - w) * (double)q.ey));
    }
};

#define SPMV32_ARGS                                                                                  \
    const Desc *__restrict__ desc, const int *__restrict__ rp, const int *__restrict__ ci,          \
        const float *__restrict__ va, const float *__restrict__ xin, float *__restrict__ partials

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_f32(SPMV32_ARGS, float* __restrict__ out) {
    spmv_body<SP_NNZ_F64, float, float, float>(desc, rp, ci, va, xin, partials, EpiStore<float>{out});
}

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_t_primal_mixed(
    SPMV32_ARGS, float* __restrict__ dx, const double* __restrict__ xa, const double* __restrict__ c,
    const double* __restrict__ lv, const double* __restrict__ uv, const double* __restrict__ atya,
    const float* __restrict__ ex, double tau, double w, double rho, float* __restrict__ dx_bar) {
    spmv_body<SP_NNZ_F64, float, float, float>(desc, rp, ci, va, xin, partials,
                                               EpiPrimalMixed{dx, xa, c, lv, uv, atya, ex, tau, w, rho, dx_bar});
}

// The same in mixed precision: 4,096 FP32 entries (16 KB, 34 KB of LDS in
// all, four workgroups per CU; 8,192 entries at three per CU are slower).
extern "C" __global__ __launch_bounds__(WG) void sp_spmv_t_primal_mixed_hot(
    SPMV32_ARGS, int ndesc, float* __restrict__ dx, const double* __restrict__ xa, const double* __restrict__ c,
    const double* __restrict__ lv, const double* __restrict__ uv, const double* __restrict__ atya,
    const float* __restrict__ ex, double tau, double w, double rho, float* __restrict__ dx_bar) {
    spmv_body_hot<SP_NNZ_F64, SP_HOT_F32, float, float, float>(
        desc, ndesc, rp, ci, va, xin, partials, EpiPrimalMixed{dx, xa, c, lv, uv, atya, ex, tau, w, rho, dx_bar});
}

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_dual_mixed(
    SPMV32_ARGS, float* __restrict__ dy, const double* __restrict__ ya, const double* __restrict__ lc,
    const double* __restrict__ uc, const double* __restrict__ axa, const float* __restrict__ ey, double sigma,
    double w, double rho) {
    spmv_body<SP_NNZ_F64, float, float, float>(desc, rp, ci, va, xin, partials,
                                               EpiDualMixed{dy, ya, lc, uc, axa, ey, sigma, w, rho});
}


extern "C" __global__ __launch_bounds__(WG) void sp_finish_f32(FIN_ARGS(float), float* __restrict__ out) {
    finish_body<float>(longs, nlong, partials, EpiStore<float>{out});
}

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_panel_f32(SPMV32_ARGS, float* __restrict__ out) {
    spmv_body<SP_NNZ_F64, float, float, float>(desc, rp, ci, va, xin, partials, EpiStore<float>{out});
}

extern "C" __global__ __launch_bounds__(WG) void sp_finish_panel_f32(FIN_ARGS(float), float* __restrict__ out) {
    finish_body<float>(longs, nlong, partials, EpiStore<float>{out});
}

extern "C" __global__ __launch_bounds__(WG) void sp_finish_t_primal_mixed(
    FIN_ARGS(float), float* __restrict__ dx, const double* __restrict__ xa, const double* __restrict__ c,
    const double* __restrict__ lv, const double* __restrict__ uv, const double* __restrict__ atya,
    const float* __restrict__ ex, double tau, double w, double rho, float* __restrict__ dx_bar) {
    finish_body<float>(longs, nlong, partials, EpiPrimalMixed{dx, xa, c, lv, uv, atya, ex, tau, w, rho, dx_bar});
}

extern "C" __global__ __launch_bounds__(WG) void sp_finish_dual_mixed(
    FIN_ARGS(float), float* __restrict__ dy, const double* __restrict__ ya, const double* __restrict__ lc,
    const double* __restrict__ uc, const double* __restrict__ axa, const float* __restrict__ ey, double sigma,
    double w, double rho) {
    finish_body<float>(longs, nlong, partials, EpiDualMixed{dy, ya, lc, uc, axa, ey, sigma, w, rho});
}

// ---- element-wise (cpu/src/lib.rs:147-209) ------------------------------------

#define GRID_LOOP(i, n) for (i64 i = (i64)blockIdx.x * WG + threadIdx.x; i < (n); i += (i64)gridDim.x * WG)

extern "C" __global__ __launch_bounds__(WG) void sp_primal_step_f64(const double* __restrict__ x,
                                                                    const double* __restrict__ aty,
                                                                    const double* __restrict__ c,
                                                                    const double* __restrict__ lv,
                                                                    const double* __restrict__ uv, double tau,
                                                                    double* __restrict__ x_t,
                                                                    double* __restrict__ x_bar, i64 n) {
    GRID_LOOP(j, n) {
        const double xj = x[j];
        const double t = clampv(xj - tau * (c[j] - aty[j]), lv[j], uv[j]);
        x_t[j] = t;
        x_bar[j] = 2.0 * t - xj;
    }
}

extern "C" __global__ __launch_bounds__(WG) void sp_dual_step_f64(const double* __restrict__ y,
                                                                  const double* __restrict__ ax,
                                                                  const double* __restrict__ lc,
                                                                  const double* __restrict__ uc, double sigma,
                                                                  double* __restrict__ y_t, i64 m) {
    GRID_LOOP(i, m) {
        const double a = ax[i], yi = y[i];
        const double v = a - yi / sigma;
        y_t[i] = yi - sigma * (a - clampv(v, lc[i], uc[i]));
    }
}

extern "C" __global__ __launch_bounds__(WG) void sp_halpern_step_f64(double* __restrict__ z,
                                                                     const double* __restrict__ z_t,
                                                                     const double* __restrict__ anchor, double w,
                                                                     double rho, i64 n) {
    GRID_LOOP(i, n) { z[i] = w * ((1.0 + rho) * z_t[i] - rho * z[i]) + (1.0 - w) * anchor[i]; }
}

extern "C" __global__ __launch_bounds__(WG) void sp_sub_f64(const double* __restrict__ a,
                                                            const double* __restrict__ b,
                                                            double* __restrict__ out, i64 n) {
    GRID_LOOP(i, n) { out[i] = a[i] - b[i]; }
}

extern "C" __global__ __launch_bounds__(WG) void sp_scale_f64(double* __restrict__ v, double alpha, i64 n) {
    GRID_LOOP(i, n) { v[i] *= alpha; }
}

// v *= 1 / sqrt(*ns): the power iteration's normalisation with the norm left
// on the device (sqrt and the division are correctly rounded, as on the host).
extern "C" __global__ __launch_bounds__(WG) void sp_scale_inv_sqrt_f64(double* __restrict__ v,
                                                                       const double* __restrict__ ns, i64 n) {
    const double alpha = 1.0 / sqrt(*ns);
    GRID_LOOP(i, n) { v[i] *= alpha; }
}

extern "C" __global__ __launch_bounds__(WG) void sp_fill_f64(double* __restrict__ v, double value, i64 n) {
    GRID_LOOP(i, n) { v[i] = value; }
}

// ---- unscaling (scaling.rs:114-121) -----------------------------------------
//
// x = x_s · col_scale / bound_scale and y = y_s · row_scale / obj_scale, the
// product first and then the (correctly rounded) division, as on the host.
// With `to`, the result lands in the LP's own order: entry q of a relabeled
// vector goes to to[q]. The primal factors are in the relabeled order (they
// were computed on the relabeled matrix), the dual factors in the LP's own.
extern "C" __global__ __launch_bounds__(WG) void sp_unscale_primal_f64(const double* __restrict__ v,
                                                                       const double* __restrict__ f, double s,
                                                                       const int* __restrict__ to,
                                                                       double* __restrict__ out, i64 n) {
    GRID_LOOP(q, n) { out[to ? to[q] : q] = v[q] * f[q] / s; }
}

extern "C" __global__ __launch_bounds__(WG) void sp_unscale_dual_f64(const double* __restrict__ v,
                                                                     const double* __restrict__ f, double s,
                                                                     const int* __restrict__ to,
                                                                     double* __restrict__ out, i64 m) {
    GRID_LOOP(i, m) {
        const i64 d = to ? to[i] : i;
        out[d] = v[i] * f[d] / s;
    }
}

// Precision conversions (round to nearest; ±inf and NaN carry over).
extern "C" __global__ __launch_bounds__(WG) void sp_f64_to_f32(const double* __restrict__ a, float* __restrict__ b,
                                                               i64 n) {
    GRID_LOOP(i, n) { b[i] = (float)a[i]; }
}

// dz = (float)(z - anchor) and back: z = anchor + dz.
extern "C" __global__ __launch_bounds__(WG) void sp_delta_from(const double* __restrict__ z,
                                                               const double* __restrict__ anchor,
                                                               float* __restrict__ dz, i64 n) {
    GRID_LOOP(i, n) { dz[i] = (float)(z[i] - anchor[i]); }
}

// z = base + dz; base may be z itself.
extern "C" __global__ __launch_bounds__(WG) void sp_delta_to(const float* __restrict__ dz, const double* anchor,
                                                             double* z, i64 n) {
    GRID_LOOP(i, n) { z[i] = anchor[i] + (double)dz[i]; }
}

// ---- fused reductions (cpu/src/lib.rs:189-259) ---------------------------------
//
// Each kernel runs over a combined index range (rows, then columns when a
// method reduces both), accumulates K sums per thread in a grid-stride loop
// and reduces them per workgroup in a fixed order. A one-workgroup launch
// writes its K totals to `result`; otherwise each workgroup writes its K
// partials and the separate finishing kernel sp_red_finish adds them in
// workgroup order (no fence or atomic: the kernel boundary orders them).
// Two launches at most, and one 8·K-byte readback per method.

// Workgroup b of a G-workgroup reduction over [0, n).
template <int K, class F>
__device__ __forceinline__ void reduce_part(i64 n, const F& f, int b, int G, double* __restrict__ partials,
                                            double* __restrict__ result) {
    double s[K];
#pragma unroll
    for (int k = 0; k < K; ++k) s[k] = 0.0;
    // Tiles of 4·WG elements; a thread takes elements tid, tid+WG, ... of a tile,
    // so its four loads are independent and issue together.
    for (i64 base = (i64)b * 4 * WG; base < n; base += (i64)G * 4 * WG) {
        if (base + 4 * WG <= n) {
#pragma unroll
            for (int u = 0; u < 4; ++u) f(base + threadIdx.x + u * WG, s);
        } else {
            for (int u = 0; u < 4; ++u) {
                const i64 i = base + threadIdx.x + u * WG;
                if (i < n) f(i, s);
            }
        }
    }
    block_sum<K>(s);
    if (threadIdx.x == 0) {
        double* out = G == 1 ? result : partials + (i64)b * K;
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = s[k];
    }
}

template <int K, class F>
__device__ __forceinline__ void reduce_body(i64 n, const F& f, double* __restrict__ partials,
                                            double* __restrict__ result) {
    reduce_part<K>(n, f, (int)blockIdx.x, (int)gridDim.x, partials, result);
}

// Adds the per-workgroup partials of a reduction (G workgroups, K sums each)
// in workgroup order: thread t takes workgroups t, t + WG, ..., then the block
// sums in its fixed order. One workgroup.
__device__ __forceinline__ void red_finish_body(const double* __restrict__ partials, int G, int K,
                                                double* __restrict__ result) {
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < G; b += WG)
#pragma unroll
        for (int k = 0; k < 5; ++k)
            if (k < K) s[k] += partials[(i64)b * K + k];
    block_sum<5>(s);
    if (threadIdx.x == 0)
        for (int k = 0; k < K; ++k) result[k] = s[k];
}

extern "C" __global__ __launch_bounds__(WG) void sp_red_finish(const double* __restrict__ partials, int G, int K,
                                                               double* __restrict__ result) {
    red_finish_body(partials, G, K, result);
}

#define RED_TAIL double *__restrict__ partials, double *__restrict__ result

struct FDot {
    const double *a, *b;
    __device__ void operator()(i64 i, double (&s)[1]) const { s[0] += a[i] * b[i]; }
};
extern "C" __global__ __launch_bounds__(WG) void sp_red_dot_f64(const double* __restrict__ a,
                                                                const double* __restrict__ b, i64 n, RED_TAIL) {
    reduce_body<1>(n, FDot{a, b}, partials, result);
}

struct FDist {
    const double *a, *b;
    __device__ void operator()(i64 i, double (&s)[1]) const {
        const double d = a[i] - b[i];
        s[0] += d * d;
    }
};
extern "C" __global__ __launch_bounds__(WG) void sp_red_distance_sq_f64(const double* __restrict__ a,
                                                                        const double* __restrict__ b, i64 n,
                                                                        RED_TAIL) {
    reduce_body<1>(n, FDist{a, b}, partials, result);
}

// primal_stats: [Σ_i (excess(ax, lc, uc) pf)², Σ_j c x]
struct FPrimal {
    const double *ax, *lc, *uc, *pf, *c, *x;
    i64 m;
    __device__ void operator()(i64 t, double (&s)[2]) const {
        if (t < m) {
            const double e = excess(ax[t], lc[t], uc[t]) * pf[t];
            s[0] += e * e;
        } else {
            const i64 j = t - m;
            s[1] += c[j] * x[j];
        }
    }
};
extern "C" __global__ __launch_bounds__(WG) void sp_red_primal_stats_f64(
    const double* __restrict__ ax, const double* __restrict__ lc, const double* __restrict__ uc,
    const double* __restrict__ pf, i64 m, const double* __restrict__ c, const double* __restrict__ x, i64 n,
    RED_TAIL) {
    reduce_body<2>(m + n, FPrimal{ax, lc, uc, pf, c, x, m}, partials, result);
}

// dual_stats: [Σ_i bound_term(y, lc, uc), Σ_j (forbidden(r) df)², Σ_j bound_term(r, lv, uv)], r = c - aty
struct FDual {
    const double *y, *lc, *uc, *c, *aty, *lv, *uv, *df;
    i64 m;
    __device__ void operator()(i64 t, double (&s)[3]) const {
        if (t < m) {
            s[0] += bound_term(y[t], lc[t], uc[t]);
        } else {
            const i64 j = t - m;
            const double r = c[j] - aty[j];
            const double f = forbidden(r, lv[j], uv[j]) * df[j];
            s[1] += f * f;
            s[2] += bound_term(r, lv[j], uv[j]);
        }
    }
};
extern "C" __global__ __launch_bounds__(WG) void sp_red_dual_stats_f64(
    const double* __restrict__ y, const double* __restrict__ lc, const double* __restrict__ uc, i64 m,
    const double* __restrict__ c, const double* __restrict__ aty, const double* __restrict__ lv,
    const double* __restrict__ uv, const double* __restrict__ df, i64 n, RED_TAIL) {
    reduce_body<3>(m + n, FDual{y, lc, uc, c, aty, lv, uv, df, m}, partials, result);
}

// step_dual_residual_sq: Σ ((aty - aty_t - (x_t - x) inv_tau) df)²
struct FStep {
    const double *x, *x_t, *aty, *aty_t, *df;
    double inv_tau;
    __device__ void operator()(i64 j, double (&s)[1]) const {
        const double r = (aty[j] - aty_t[j] - (x_t[j] - x[j]) * inv_tau) * df[j];
        s[0] += r * r;
    }
};
extern "C" __global__ __launch_bounds__(WG) void sp_red_step_dual_f64(
    const double* __restrict__ x, const double* __restrict__ x_t, const double* __restrict__ aty,
    const double* __restrict__ aty_t, const double* __restrict__ df, double inv_tau, i64 n, RED_TAIL) {
    reduce_body<1>(n, FStep{x, x_t, aty, aty_t, df, inv_tau}, partials, result);
}

// primal_ray_stats: [Σ_i rv(adx, lc, uc)², Σ_j rv(dx, lv, uv)², Σ_j dx², Σ_j c dx]
struct FPrimalRay {
    const double *adx, *lc, *uc, *dx, *c, *lv, *uv;
    i64 m;
    __device__ void operator()(i64 t, double (&s)[4]) const {
        if (t < m) {
            const double v = recession_violation(adx[t], lc[t], uc[t]);
            s[0] += v * v;
        } else {
            const i64 j = t - m;
            const double d = dx[j];
            const double v = recession_violation(d, lv[j], uv[j]);
            s[1] += v * v;
            s[2] += d * d;
            s[3] += c[j] * d;
        }
    }
};
extern "C" __global__ __launch_bounds__(WG) void sp_red_primal_ray_f64(
    const double* __restrict__ adx, const double* __restrict__ lc, const double* __restrict__ uc, i64 m,
    const double* __restrict__ dx, const double* __restrict__ c, const double* __restrict__ lv,
    const double* __restrict__ uv, i64 n, RED_TAIL) {
    reduce_body<4>(m + n, FPrimalRay{adx, lc, uc, dx, c, lv, uv, m}, partials, result);
}

// dual_ray_stats: [Σ_i forbidden(dy)², Σ_i bound_term(dy), Σ_i dy², Σ_j forbidden(r)², Σ_j bound_term(r)], r = -atdy
struct FDualRay {
    const double *dy, *lc, *uc, *atdy, *lv, *uv;
    i64 m;
    __device__ void operator()(i64 t, double (&s)[5]) const {
        if (t < m) {
            const double d = dy[t];
            const double f = forbidden(d, lc[t], uc[t]);
            s[0] += f * f;
            s[1] += bound_term(d, lc[t], uc[t]);
            s[2] += d * d;
        } else {
            const i64 j = t - m;
            const double r = -atdy[j];
            const double f = forbidden(r, lv[j], uv[j]);
            s[3] += f * f;
            s[4] += bound_term(r, lv[j], uv[j]);
        }
    }
};
extern "C" __global__ __launch_bounds__(WG) void sp_red_dual_ray_f64(
    const double* __restrict__ dy, const double* __restrict__ lc, const double* __restrict__ uc, i64 m,
    const double* __restrict__ atdy, const double* __restrict__ lv, const double* __restrict__ uv, i64 n,
    RED_TAIL) {
    reduce_body<5>(m + n, FDualRay{dy, lc, uc, atdy, lv, uv, m}, partials, result);
}

// ---- preconditioning passes (scaling.rs:53-103, 139-184) ---------------------------
//
// The scaling loop runs on the device over A and its transpose A'ᵀ (whose rows
// are A's columns, in A's row order). Every pass reproduces the host's
// arithmetic: row and column extrema are exact whatever the order; the
// Pock-Chambolle sums add a line's terms left to right as the host does (for
// a column, in row order, which is A'ᵀ's order); a factor is applied as
// v·(r_i·c_j) to both copies (IEEE products commute), so both stay equal to
// the host's scaled matrix and its transpose.

__device__ __forceinline__ double inv_sqrt_or_one(double v) {
    return (v > 0.0 && __builtin_isfinite(v)) ? 1.0 / sqrt(v) : 1.0;
}

// Per row (one wave per row): lo = min |a|, hi = max |a| (lo = 0 for a row
// without nonzero magnitude, scaling.rs:159-178). lo may be null (max only).
extern "C" __global__ __launch_bounds__(WG) void sp_sc_minmax(const int* __restrict__ rp,
                                                              const double* __restrict__ va, int nrows,
                                                              double* __restrict__ lo, double* __restrict__ hi) {
    const int row = (int)((blockIdx.x * WG + threadIdx.x) / WAVE), lane = threadIdx.x % WAVE;
    if (row >= nrows) return;  // uniform per wave
    double mn = __builtin_inf(), mx = 0.0;
    for (int k = rp[row] + lane; k < rp[row + 1]; k += WAVE) {
        const double v = fabs(va[k]);
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const double a = __shfl_xor(mn, o, WAVE), b = __shfl_xor(mx, o, WAVE);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    if (lane == 0) {
        hi[row] = mx;
        if (lo) lo[row] = mx == 0.0 ? 0.0 : mn;
    }
}

// Per row (one thread per row, left to right as the host): Σ |a|^e.
extern "C" __global__ __launch_bounds__(WG) void sp_sc_powsum(const int* __restrict__ rp,
                                                              const double* __restrict__ va, int nrows, double e,
                                                              double* __restrict__ out) {
    const int row = blockIdx.x * WG + threadIdx.x;
    if (row >= nrows) return;
    double s = 0.0;
    const int k0 = rp[row], k1 = rp[row + 1];
    if (e == 1.0) {
        for (int k = k0; k < k1; ++k) s = s + fabs(va[k]);
    } else {
        for (int k = k0; k < k1; ++k) s = s + pow(fabs(va[k]), e);
    }
    out[row] = s;
}

// f = inv_sqrt(a·b) (geometric mean, b non-null) or inv_sqrt(a).
extern "C" __global__ __launch_bounds__(WG) void sp_sc_factor(const double* __restrict__ a,
                                                              const double* __restrict__ b,
                                                              double* __restrict__ f, i64 n) {
    GRID_LOOP(i, n) { f[i] = inv_sqrt_or_one(b ? a[i] * b[i] : a[i]); }
}

// v_k ← v_k·(rf[row]·cf[col_k]) over a CSR matrix (one wave per row); a null
// factor vector reads as ones.
extern "C" __global__ __launch_bounds__(WG) void sp_sc_apply(const int* __restrict__ rp, const int* __restrict__ ci,
                                                             double* __restrict__ va, int nrows,
                                                             const double* __restrict__ rf,
                                                             const double* __restrict__ cf) {
    const int row = (int)((blockIdx.x * WG + threadIdx.x) / WAVE), lane = threadIdx.x % WAVE;
    if (row >= nrows) return;
    const double r = rf ? rf[row] : 1.0;
    for (int k = rp[row] + lane; k < rp[row + 1]; k += WAVE) va[k] = va[k] * (r * (cf ? cf[ci[k]] : 1.0));
}

// v ← v·f element-wise (the accumulated row and column scales).
extern "C" __global__ __launch_bounds__(WG) void sp_sc_mul(double* __restrict__ v, const double* __restrict__ f,
                                                           i64 n) {
    GRID_LOOP(i, n) { v[i] *= f[i]; }
}

// ---- transpose on the device (problem.rs:88-109) ---------------------------------------
//
// A'ᵀ is A's nonzeros ordered by column, and within a column by row: a stable
// sort of the nonzero indices k by column index (A's nonzeros are already in
// row order). The sort is a least-significant-digit radix sort, 8 bits per
// pass: per-tile digit counts, an exclusive scan of the counts (digit-major,
// tile-minor), then a stable scatter in which each item's rank among equal
// digits of its tile comes from wave ballots. Everything is integer work, so
// the result is the host's counting sort exactly.

#define TR_ITEMS 16                 // items per thread per tile
#define TR_TILE (WG * TR_ITEMS)     // 4096 nonzeros per workgroup

// rows[k] = the row of nonzero k (one wave per row).
extern "C" __global__ __launch_bounds__(WG) void sp_tr_rows(const int* __restrict__ rp, int nrows,
                                                            int* __restrict__ rows) {
    const int row = (int)((blockIdx.x * WG + threadIdx.x) / WAVE), lane = threadIdx.x % WAVE;
    if (row >= nrows) return;
    for (int k = rp[row] + lane; k < rp[row + 1]; k += WAVE) rows[k] = row;
}

// hist[d·ntiles + tile] = items of the tile whose digit is d.
extern "C" __global__ __launch_bounds__(WG) void sp_tr_hist(const int* __restrict__ keys, int nnz, int shift,
                                                            int* __restrict__ hist, int ntiles) {
    __shared__ int cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const i64 base = (i64)blockIdx.x * TR_TILE;
    for (int c = 0; c < TR_ITEMS; ++c) {
        const i64 idx = base + c * WG + threadIdx.x;
        if (idx < nnz) atomicAdd(&cnt[(keys[idx] >> shift) & 255], 1);
    }
    __syncthreads();
    hist[(i64)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan, level 1: each workgroup scans 4·WG ints in place and
// writes its total to sums[block].
extern "C" __global__ __launch_bounds__(WG) void sp_scan_tiles(int* __restrict__ a, i64 len,
                                                               int* __restrict__ sums) {
    __shared__ int wtot[NWAVES];
    const i64 base = (i64)blockIdx.x * 4 * WG + 4 * threadIdx.x;
    int v[4], t = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        v[u] = base + u < len ? a[base + u] : 0;
        t += v[u];
    }
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    int incl = t;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int y = __shfl_up(incl, o, WAVE);
        if (lane >= o) incl += y;
    }
    if (lane == WAVE - 1) wtot[wave] = incl;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; ++w) before += wtot[w];
    int run = before + incl - t;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (base + u < len) a[base + u] = run;
        run += v[u];
    }
    if (threadIdx.x == WG - 1) sums[blockIdx.x] = run;
}

// Level 2: a[i] += offsets[block of i].
extern "C" __global__ __launch_bounds__(WG) void sp_scan_add(int* __restrict__ a, i64 len,
                                                             const int* __restrict__ offsets) {
    const int add = offsets[blockIdx.x];
    const i64 base = (i64)blockIdx.x * 4 * WG;
    for (int u = 0; u < 4; ++u) {
        const i64 i = base + u * WG + threadIdx.x;
        if (i < len) a[i] += add;
    }
}

// The stable scatter of one pass: the item at position idx of the tile goes to
// offsets[d·ntiles + tile] + (items of digit d earlier in the tile).
extern "C" __global__ __launch_bounds__(WG) void sp_tr_scatter(const int* __restrict__ keys_in,
                                                               const int* __restrict__ vals_in, int nnz, int shift,
                                                               const int* __restrict__ offsets, int ntiles,
                                                               int* __restrict__ keys_out,
                                                               int* __restrict__ vals_out) {
    __shared__ int run[256];
    __shared__ int base[256];
    __shared__ int wcount[NWAVES][256];
    const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    run[tid] = 0;
    base[tid] = offsets[(i64)tid * ntiles + blockIdx.x];
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int c = 0; c < TR_ITEMS; ++c) {
#pragma unroll
        for (int w = 0; w < NWAVES; ++w) wcount[w][tid] = 0;
        __syncthreads();
        const i64 idx = (i64)blockIdx.x * TR_TILE + c * WG + tid;
        const bool valid = idx < nnz;
        const int key = valid ? keys_in[idx] : 0;
        const int val = valid ? vals_in[idx] : 0;
        const int d = (key >> shift) & 255;
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const unsigned long long on = __ballot(valid && ((d >> bit) & 1));
            m &= ((d >> bit) & 1) ? on : ~on;
        }
        const int rank = __popcll(m & below);
        if (valid && rank == 0) wcount[wave][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            int pos = base[d] + run[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcount[w][d];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int w = 0; w < NWAVES; ++w) add += wcount[w][tid];
        run[tid] += add;
    }
}

// Nonzeros per column (integer counts: the same in any order).
extern "C" __global__ __launch_bounds__(WG) void sp_col_count(const int* __restrict__ ci, i64 n, int* __restrict__ cnt) {
    GRID_LOOP(k, n) { atomicAdd(cnt + ci[k], 1); }
}

// ci[k] ← rank[ci[k]] (the column relabeling of the driver, done on the device).
extern "C" __global__ __launch_bounds__(WG) void sp_remap(int* __restrict__ ci, const int* __restrict__ rank, i64 n) {
    GRID_LOOP(k, n) { ci[k] = rank[ci[k]]; }
}

extern "C" __global__ __launch_bounds__(WG) void sp_tr_iota(int* __restrict__ a, i64 n) {
    GRID_LOOP(i, n) { a[i] = (int)i; }
}

// Column pointers from the sorted column indices: tp[j] = first position of
// column j (empty columns included), tp[ncols] = nnz.
extern "C" __global__ __launch_bounds__(WG) void sp_tr_ptr(const int* __restrict__ keys, int nnz, int ncols,
                                                           int* __restrict__ tp) {
    const i64 q = (i64)blockIdx.x * WG + threadIdx.x;
    if (q > nnz) return;
    const int lo = q == 0 ? 0 : keys[q - 1] + 1;
    const int hi = q == nnz ? ncols : keys[q];
    for (int j = lo; j <= hi; ++j) tp[j] = (int)q;
}

// A'ᵀ's row indices and values from the sorted permutation.
extern "C" __global__ __launch_bounds__(WG) void sp_tr_gather(const int* __restrict__ perm,
                                                              const int* __restrict__ rows,
                                                              const double* __restrict__ va, i64 nnz,
                                                              int* __restrict__ tci, double* __restrict__ tva) {
    GRID_LOOP(q, nnz) {
        const int k = perm[q];
        tci[q] = rows[k];
        tva[q] = va[k];
    }
}

// ---- the KKT checker (crates/saddlepoint-check/src/lib.rs:62-118) ---------------------
//
// One pass over rows then columns of a claimed (x, y) with Ax and Aᵀy of the
// original matrix: seven sums, [Σ c·x, Σ outside(Ax, lc, uc)², Σ outside(x,
// lv, uv)², dual objective terms, Σ dual-sign violations², ‖b‖², ‖c‖²].

__device__ __forceinline__ double outside(double v, double lo, double hi) {
    return v < lo ? lo - v : (v > hi ? v - hi : 0.0);
}

// check/src/lib.rs:84-90: a positive multiplier needs a finite lower bound,
// a negative one a finite upper bound; otherwise it is a violation.
__device__ __forceinline__ void dual_term(double v, double lo, double hi, double& obj, double& viol) {
    if (v > 0.0) {
        if (__builtin_isfinite(lo)) obj += v * lo;
        else viol += v * v;
    } else if (v < 0.0) {
        if (__builtin_isfinite(hi)) obj += v * hi;
        else viol += v * v;
    }
}

struct FCheck {
    const double *ax, *y, *lc, *uc, *x, *aty, *c, *lv, *uv;
    i64 m;
    __device__ void operator()(i64 t, double (&s)[7]) const {
        if (t < m) {
            const double o = outside(ax[t], lc[t], uc[t]);
            s[1] += o * o;
            dual_term(y[t], lc[t], uc[t], s[3], s[4]);
            const double l = __builtin_isfinite(lc[t]) ? lc[t] : 0.0;
            const double u = (__builtin_isfinite(uc[t]) && uc[t] != lc[t]) ? uc[t] : 0.0;
            s[5] += l * l + u * u;
        } else {
            const i64 j = t - m;
            s[0] += c[j] * x[j];
            const double o = outside(x[j], lv[j], uv[j]);
            s[2] += o * o;
            dual_term(c[j] - aty[j], lv[j], uv[j], s[3], s[4]);
            s[6] += c[j] * c[j];
        }
    }
};

extern "C" __global__ __launch_bounds__(WG) void sp_red_check_f64(
    const double* __restrict__ ax, const double* __restrict__ y, const double* __restrict__ lc,
    const double* __restrict__ uc, i64 m, const double* __restrict__ x, const double* __restrict__ aty,
    const double* __restrict__ c, const double* __restrict__ lv, const double* __restrict__ uv, i64 n, RED_TAIL) {
    reduce_body<7>(m + n, FCheck{ax, y, lc, uc, x, aty, c, lv, uv, m}, partials, result);
}

// The per-workgroup partials of sp_red_check_f64, in workgroup order.
extern "C" __global__ __launch_bounds__(WG) void sp_red_finish7(const double* __restrict__ partials, int G,
                                                                double* __restrict__ result) {
    double s[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < G; b += WG)
#pragma unroll
        for (int k = 0; k < 7; ++k) s[k] += partials[(i64)b * 7 + k];
    block_sum<7>(s);
    if (threadIdx.x == 0)
        for (int k = 0; k < 7; ++k) result[k] = s[k];
}

// ---- bandwidth probes (the counterpart of the reference's sp-bandwidth) ------------
//
// Achievable HBM copy and triad rates, and the rate of random 8-byte gathers
// from a table of a given size, to put the SpMV's figures in context.

extern "C" __global__ __launch_bounds__(WG) void sp_probe_copy(const double* __restrict__ a, double* __restrict__ b,
                                                               i64 n) {
    GRID_LOOP(i, n) { b[i] = __builtin_nontemporal_load(a + i); }
}

extern "C" __global__ __launch_bounds__(WG) void sp_probe_triad(const double* __restrict__ a,
                                                                const double* __restrict__ b, double* __restrict__ c,
                                                                double s, i64 n) {
    GRID_LOOP(i, n) { c[i] = __builtin_nontemporal_load(a + i) + s * __builtin_nontemporal_load(b + i); }
}

// out[i] = table[idx[i]]: streams idx and out, gathers from table.
// Eight independent gathers in flight per thread, as in the SpMV's stream
// blocks, so the rate measured is the memory system's, not one load's latency.
extern "C" __global__ __launch_bounds__(WG) void sp_probe_gather(const int* __restrict__ idx,
                                                                 const double* __restrict__ table,
                                                                 double* __restrict__ out, i64 n) {
    const i64 stride = (i64)gridDim.x * WG;
    for (i64 base = (i64)blockIdx.x * WG + threadIdx.x; base < n; base += 8 * stride) {
        int c[8];
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const i64 i = base + u * stride;
            c[u] = i < n ? __builtin_nontemporal_load(idx + i) : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = table[c[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const i64 i = base + u * stride;
            if (i < n) out[i] = v[u];
        }
    }
}

// Gathers beside a stream: blocks of WG·8 elements. A gather block loads
// 2048 indices (coalesced) and gathers table[idx] for each; a stream block
// loads 2048 doubles and 2048 ints (12 bytes an element, like a matrix's
// values and column indices). mode 0: gather blocks only; 1: stream blocks
// only; 2: the two kinds in alternate workgroups (both over all nb blocks);
// 3: every workgroup a stream block, then a gather block (as an SpMV stream
// block does). Each thread stores one sum (never read), so nothing is elided.
__device__ __forceinline__ double probe_gather_block(i64 b, const int* __restrict__ idx,
                                                     const double* __restrict__ table) {
    const i64 base = b * WG * 8;
    int c[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) c[u] = idx[base + threadIdx.x + u * WG];
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) s += table[c[u]];
    return s;
}
__device__ __forceinline__ double probe_stream_block(i64 b, const double* __restrict__ v,
                                                     const int* __restrict__ ci) {
    const i64 base = b * WG * 8;
    double x[8];
    int y[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        x[u] = v[base + threadIdx.x + u * WG];
        y[u] = ci[base + threadIdx.x + u * WG];
    }
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) s += x[u] + (double)y[u];
    return s;
}
extern "C" __global__ __launch_bounds__(WG) void sp_probe_overlap(const int* __restrict__ idx,
                                                                  const double* __restrict__ table,
                                                                  const double* __restrict__ v,
                                                                  const int* __restrict__ ci, int mode,
                                                                  double* __restrict__ out) {
    const i64 b = blockIdx.x;
    double s;
    if (mode == 0) s = probe_gather_block(b, idx, table);
    else if (mode == 1) s = probe_stream_block(b, v, ci);
    else if (mode == 2) s = (b & 1) ? probe_stream_block(b >> 1, v, ci) : probe_gather_block(b >> 1, idx, table);
    else s = probe_stream_block(b, v, ci) + probe_gather_block(b, idx, table);
    out[b * WG + threadIdx.x] = s;
}

extern "C" __global__ __launch_bounds__(WG) void sp_probe_copy_f32(const float* __restrict__ a, float* __restrict__ b,
                                                                   i64 n) {
    GRID_LOOP(i, n) { b[i] = __builtin_nontemporal_load(a + i); }
}

extern "C" __global__ __launch_bounds__(WG) void sp_probe_triad_f32(const float* __restrict__ a,
                                                                    const float* __restrict__ b, float* __restrict__ c,
                                                                    float s, i64 n) {
    GRID_LOOP(i, n) { c[i] = __builtin_nontemporal_load(a + i) + s * __builtin_nontemporal_load(b + i); }
}

// The gathers of an SpMV and nothing else: workgroup b takes work item
// desc[b] of a matrix's SpMV plan (a stream block's nonzeros [k0, k1), or a
// long-row chunk's) and, in batches of SP_NNZ_F64 with the stream blocks'
// straight-line loads (the same lanes, the same order, all in flight), loads
// each nonzero's column index and table[col]. Values, row pointers, the
// reduction and the epilogue are left out, so its rate is the ceiling that
// this matrix's access pattern puts on the SpMV's gathers. The sum is stored
// only if it is NaN, which a table of finite values never gives.
template <class T>
__device__ __forceinline__ void replay_body(const Desc* __restrict__ desc, const int* __restrict__ ci,
                                            const T* __restrict__ table, double* __restrict__ sink) {
    constexpr int PER = SP_NNZ_F64 / WG;
    const Desc d = desc[blockIdx.x];
    const int tid = threadIdx.x;
    double s = 0.0;
    for (int base = d.k0; base < d.k1; base += SP_NNZ_F64) {
        const int n = min(SP_NNZ_F64, d.k1 - base), last = n - 1;
        int c[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) c[u] = ci[base + min(tid + u * WG, last)];
        T v[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) v[u] = table[c[u]];
#pragma unroll
        for (int u = 0; u < PER; ++u)
            if (tid + u * WG < n) s += (double)v[u];
    }
    if (s != s) sink[0] = s;
}

extern "C" __global__ __launch_bounds__(WG) void sp_probe_replay_f64(const Desc* __restrict__ desc,
                                                                     const int* __restrict__ ci,
                                                                     const double* __restrict__ table,
                                                                     double* __restrict__ sink) {
    replay_body<double>(desc, ci, table, sink);
}

extern "C" __global__ __launch_bounds__(WG) void sp_probe_replay_f32(const Desc* __restrict__ desc,
                                                                     const int* __restrict__ ci,
                                                                     const float* __restrict__ table,
                                                                     double* __restrict__ sink) {
    replay_body<float>(desc, ci, table, sink);
}

// ---- batched plain iterations for the serving loop --------------------------------
//
// One launch runs the plain iteration of every LP in the batch: the work items
// of all LPs' own SpMV plans are concatenated, each tagged with its LP, and a
// table holds each LP's arrays and step parameters. Every row is computed by
// the same code, in the same order, as when its LP runs alone. The Halpern
// weight w = (k+1)/(k+2) is formed on the device from the LP's iteration
// counter at the chunk start plus the step index, so a captured graph of the
// steps stays valid from chunk to chunk.

struct BInst {
    // A' (rows = m) and A'ᵀ (rows = n): row pointers, column indices, FP64 and
    // FP32 values, long-row chunk sums.
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
    float *dx, *dy, *dx_bar;       // mixed precision: displacements from the base (x, y),
    const double *atya, *axa;      //   the base's FP64 products,
    const float *ex, *ey;          //   anchor - base
    double tau, sigma, rho;
    unsigned long long k0;         // Halpern counter k at step 0 of the chunk
};

static_assert(sizeof(BInst) == 248, "BInst layout is shared with the host");

struct BDesc {
    Desc d;
    int inst, pad0, pad1, pad2;
};

struct BLong {
    int inst, row, pbase, nch;
};

__device__ __forceinline__ double halpern_w(unsigned long long k0, int it) {
    const unsigned long long k = k0 + (unsigned long long)it;
    return (double)(k + 1) / (double)(k + 2);
}

extern "C" __global__ __launch_bounds__(WG) void sp_bspmv_t_primal_halpern_f64(const BDesc* __restrict__ desc,
                                                                               const BInst* __restrict__ insts,
                                                                               int it) {
    const BDesc bd = desc[blockIdx.x];
    const BInst& I = insts[bd.inst];
    spmv_block<SP_NNZ_F64, double, double, double>(
        bd.d, I.t_rp, I.t_ci, I.t_va, (const double*)I.y, I.t_part,
        EpiPrimalHalpern{I.x, I.c, I.lv, I.uv, I.x_anchor, I.tau, halpern_w(I.k0, it), I.rho, I.x_bar});
}

extern "C" __global__ __launch_bounds__(WG) void sp_bspmv_dual_halpern_f64(const BDesc* __restrict__ desc,
                                                                           const BInst* __restrict__ insts, int it) {
    const BDesc bd = desc[blockIdx.x];
    const BInst& I = insts[bd.inst];
    spmv_block<SP_NNZ_F64, double, double, double>(
        bd.d, I.a_rp, I.a_ci, I.a_va, (const double*)I.x_bar, I.a_part,
        EpiDualHalpern{I.y, I.lc, I.uc, I.y_anchor, I.sigma, halpern_w(I.k0, it), I.rho});
}

extern "C" __global__ __launch_bounds__(WG) void sp_bfinish_t_primal_halpern_f64(const BLong* __restrict__ longs,
                                                                                 int nlong,
                                                                                 const BInst* __restrict__ insts,
                                                                                 int it) {
    const int w = (int)((blockIdx.x * WG + threadIdx.x) / WAVE);
    if (w >= nlong) return;
    const BLong L = longs[w];
    const BInst& I = insts[L.inst];
    const LongRow lr{L.row, L.pbase, L.nch, 0};
    finish_body<double>(&lr, 1, I.t_part,
                        EpiPrimalHalpern{I.x, I.c, I.lv, I.uv, I.x_anchor, I.tau, halpern_w(I.k0, it), I.rho, I.x_bar},
                        0);
}

extern "C" __global__ __launch_bounds__(WG) void sp_bfinish_dual_halpern_f64(const BLong* __restrict__ longs, int nlong,
                                                                             const BInst* __restrict__ insts, int it) {
    const int w = (int)((blockIdx.x * WG + threadIdx.x) / WAVE);
    if (w >= nlong) return;
    const BLong L = longs[w];
    const BInst& I = insts[L.inst];
    const LongRow lr{L.row, L.pbase, L.nch, 0};
    finish_body<double>(&lr, 1, I.a_part, EpiDualHalpern{I.y, I.lc, I.uc, I.y_anchor, I.sigma, halpern_w(I.k0, it), I.rho},
                        0);
}

extern "C" __global__ __launch_bounds__(WG) void sp_bspmv_t_primal_mixed(const BDesc* __restrict__ desc,
                                                                         const BInst* __restrict__ insts, int it) {
    const BDesc bd = desc[blockIdx.x];
    const BInst& I = insts[bd.inst];
    spmv_block<SP_NNZ_F64, float, float, float>(
        bd.d, I.t_rp, I.t_ci, I.t_va32, (const float*)I.dy, (float*)I.t_part,
        EpiPrimalMixed{I.dx, I.x, I.c, I.lv, I.uv, I.atya, I.ex, I.tau, halpern_w(I.k0, it), I.rho, I.dx_bar});
}

extern "C" __global__ __launch_bounds__(WG) void sp_bspmv_dual_mixed(const BDesc* __restrict__ desc,
                                                                     const BInst* __restrict__ insts, int it) {
    const BDesc bd = desc[blockIdx.x];
    const BInst& I = insts[bd.inst];
    spmv_block<SP_NNZ_F64, float, float, float>(
        bd.d, I.a_rp, I.a_ci, I.a_va32, (const float*)I.dx_bar, (float*)I.a_part,
        EpiDualMixed{I.dy, I.y, I.lc, I.uc, I.axa, I.ey, I.sigma, halpern_w(I.k0, it), I.rho});
}

extern "C" __global__ __launch_bounds__(WG) void sp_bfinish_t_primal_mixed(const BLong* __restrict__ longs, int nlong,
                                                                           const BInst* __restrict__ insts, int it) {
    const int w = (int)((blockIdx.x * WG + threadIdx.x) / WAVE);
    if (w >= nlong) return;
    const BLong L = longs[w];
    const BInst& I = insts[L.inst];
    const LongRow lr{L.row, L.pbase, L.nch, 0};
    finish_body<float>(&lr, 1, (const float*)I.t_part,
                       EpiPrimalMixed{I.dx, I.x, I.c, I.lv, I.uv, I.atya, I.ex, I.tau, halpern_w(I.k0, it), I.rho,
                                      I.dx_bar},
                       0);
}

extern "C" __global__ __launch_bounds__(WG) void sp_bfinish_dual_mixed(const BLong* __restrict__ longs, int nlong,
                                                                       const BInst* __restrict__ insts, int it) {
    const int w = (int)((blockIdx.x * WG + threadIdx.x) / WAVE);
    if (w >= nlong) return;
    const BLong L = longs[w];
    const BInst& I = insts[L.inst];
    const LongRow lr{L.row, L.pbase, L.nch, 0};
    finish_body<float>(&lr, 1, (const float*)I.a_part,
                       EpiDualMixed{I.dy, I.y, I.lc, I.uc, I.axa, I.ey, I.sigma, halpern_w(I.k0, it), I.rho}, 0);
}

// ---- batched check iterations for the serving loop --------------------------------
//
// A check iteration (and the first iteration of a restart epoch) is a fixed
// sequence of SpMVs, element-wise passes and reductions. The serving loop runs
// each operation of that sequence once for all LPs that are due: a launch
// covers every LP's work items, each tagged with the LP's entry in an operand
// table. An LP's share of a launch is the computation of the single-LP kernel
// with the same partition and the same order (SpMV: the LP's own plan;
// reductions: the LP's own workgroup count, partials added by the same
// finishing code), so its results are bitwise those of a single solve.

struct COp {
    const int *rp, *ci;      // SpMV: the matrix (A' or A'ᵀ) and its
    const double* va;        //   long-row chunk sums
    double* part;
    const double* in;        // SpMV input
    double *o0, *o1, *o2;    // outputs
    const double *a0, *a1, *a2, *a3, *a4, *a5, *a6, *a7;  // operands
    float* f0;               // FP32 displacement (mixed precision)
    double s0, s1;           // scalars
    i64 m, n;                // lengths (element-wise ops use n)
    double* dst;             // reductions: where the K sums go
    int G, pbase;            // reductions: workgroups, offset of its partials
    int kind, K;             // sp_cred_any_f64: which reduction, its sum count
};

static_assert(sizeof(COp) == 192, "COp layout is shared with the host");

// A workgroup of an element-wise pass or a reduction: entry `op`, tile or
// workgroup index `b`.
struct CBlk {
    int op, b;
};

#define CTILE (4 * WG)

// ---- SpMVs ----

extern "C" __global__ __launch_bounds__(WG) void sp_cspmv_store_f64(const BDesc* __restrict__ desc,
                                                                    const COp* __restrict__ ops) {
    const BDesc bd = desc[blockIdx.x];
    const COp& o = ops[bd.inst];
    spmv_block<SP_NNZ_F64, double, double, double>(bd.d, o.rp, o.ci, o.va, o.in, o.part, EpiStore<double>{o.o0});
}

// spmv_t + primal_step: in = y; a0..a3 = x, c, lv, uv; s0 = τ; o0..o2 = aty, x_t, x_bar.
extern "C" __global__ __launch_bounds__(WG) void sp_cspmv_primal_f64(const BDesc* __restrict__ desc,
                                                                     const COp* __restrict__ ops) {
    const BDesc bd = desc[blockIdx.x];
    const COp& o = ops[bd.inst];
    spmv_block<SP_NNZ_F64, double, double, double>(bd.d, o.rp, o.ci, o.va, o.in, o.part,
                                                   EpiPrimal{o.a0, o.a1, o.a2, o.a3, o.s0, o.o0, o.o1, o.o2});
}

// spmv + dual_step: in = x_bar; a0..a2 = y, lc, uc; s0 = σ; o0, o1 = ax, y_t.
extern "C" __global__ __launch_bounds__(WG) void sp_cspmv_dual_f64(const BDesc* __restrict__ desc,
                                                                   const COp* __restrict__ ops) {
    const BDesc bd = desc[blockIdx.x];
    const COp& o = ops[bd.inst];
    spmv_block<SP_NNZ_F64, double, double, double>(bd.d, o.rp, o.ci, o.va, o.in, o.part,
                                                   EpiDual{o.a0, o.a1, o.a2, o.s0, o.o0, o.o1});
}

#define CFIN_HEAD                                                  \
    const int w = (int)((blockIdx.x * WG + threadIdx.x) / WAVE);  \
    if (w >= nlong) return;                                        \
    const BLong L = longs[w];                                      \
    const COp& o = ops[L.inst];                                    \
    const LongRow lr{L.row, L.pbase, L.nch, 0};

extern "C" __global__ __launch_bounds__(WG) void sp_cfinish_store_f64(const BLong* __restrict__ longs, int nlong,
                                                                      const COp* __restrict__ ops) {
    CFIN_HEAD
    finish_body<double>(&lr, 1, o.part, EpiStore<double>{o.o0}, 0);
}

extern "C" __global__ __launch_bounds__(WG) void sp_cfinish_primal_f64(const BLong* __restrict__ longs, int nlong,
                                                                       const COp* __restrict__ ops) {
    CFIN_HEAD
    finish_body<double>(&lr, 1, o.part, EpiPrimal{o.a0, o.a1, o.a2, o.a3, o.s0, o.o0, o.o1, o.o2}, 0);
}

extern "C" __global__ __launch_bounds__(WG) void sp_cfinish_dual_f64(const BLong* __restrict__ longs, int nlong,
                                                                     const COp* __restrict__ ops) {
    CFIN_HEAD
    finish_body<double>(&lr, 1, o.part, EpiDual{o.a0, o.a1, o.a2, o.s0, o.o0, o.o1}, 0);
}

// ---- element-wise passes: one CTILE-element tile per workgroup ----

#define CELEM_HEAD                                              \
    const CBlk k = blk[blockIdx.x];                             \
    const COp& o = ops[k.op];                                   \
    const i64 i0 = (i64)k.b * CTILE, i1 = min(i0 + CTILE, o.n);

// o0 = a0 - a1
extern "C" __global__ __launch_bounds__(WG) void sp_csub_f64(const CBlk* __restrict__ blk,
                                                             const COp* __restrict__ ops) {
    CELEM_HEAD
    for (i64 i = i0 + threadIdx.x; i < i1; i += WG) o.o0[i] = o.a0[i] - o.a1[i];
}

// o0 = z ← s0((1+s1) a0 - s1 z) + (1-s0) a1   (z_t = a0, anchor = a1, w = s0, ρ = s1)
extern "C" __global__ __launch_bounds__(WG) void sp_chalpern_f64(const CBlk* __restrict__ blk,
                                                                 const COp* __restrict__ ops) {
    CELEM_HEAD
    const double w = o.s0, rho = o.s1;
    for (i64 i = i0 + threadIdx.x; i < i1; i += WG)
        o.o0[i] = w * ((1.0 + rho) * o.a0[i] - rho * o.o0[i]) + (1.0 - w) * o.a1[i];
}

// o0 = a0 + f0 (base + FP32 displacement; a0 may be o0)
extern "C" __global__ __launch_bounds__(WG) void sp_cdelta_to(const CBlk* __restrict__ blk,
                                                              const COp* __restrict__ ops) {
    CELEM_HEAD
    for (i64 i = i0 + threadIdx.x; i < i1; i += WG) o.o0[i] = o.a0[i] + (double)o.f0[i];
}

// f0 = (float)(a0 - a1)
extern "C" __global__ __launch_bounds__(WG) void sp_cdelta_from(const CBlk* __restrict__ blk,
                                                                const COp* __restrict__ ops) {
    CELEM_HEAD
    for (i64 i = i0 + threadIdx.x; i < i1; i += WG) o.f0[i] = (float)(o.a0[i] - o.a1[i]);
}

// o0 *= 1 / sqrt(a0[0]) (the power iteration's normalisation, a0 a result slot)
extern "C" __global__ __launch_bounds__(WG) void sp_cscale_inv_sqrt_f64(const CBlk* __restrict__ blk,
                                                                        const COp* __restrict__ ops) {
    CELEM_HEAD
    const double alpha = 1.0 / sqrt(o.a0[0]);
    for (i64 i = i0 + threadIdx.x; i < i1; i += WG) o.o0[i] *= alpha;
}

// o0 = a0 (gathers every LP's result slots into one buffer for one readback,
// and the restart copies)
extern "C" __global__ __launch_bounds__(WG) void sp_ccopy_f64(const CBlk* __restrict__ blk,
                                                              const COp* __restrict__ ops) {
    CELEM_HEAD
    for (i64 i = i0 + threadIdx.x; i < i1; i += WG) o.o0[i] = o.a0[i];
}

// ---- reductions of several kinds in one launch ----
//
// The reductions of a check iteration (and of a power-iteration round) read
// no output of one another, so they run as one launch and one finishing
// pass over a table of entries: entry i's workgroups b < G compute exactly
// what its single-method launch computes (same workgroup count G, same
// order, same finishing code), so every result is bitwise that of the
// method. Used by the single solve (one table per solver) and the serving
// loop (every due LP's entries). Kinds, with their operands and lengths:
//   RK_DOT     a0·a1 over n
//   RK_DIST    ‖a0 - a1‖² over n
//   RK_PRIMAL  primal_stats: a0..a5 = ax, lc, uc, pf, c, x over m + n (2 sums)
//   RK_DUAL    dual_stats: a0..a7 = y, lc, uc, c, aty, lv, uv, df over m + n (3)
//   RK_STEP    step_dual_residual_sq: a0..a4 = x, x_t, aty, aty_t, df; s0 = 1/τ; over n
//   RK_PRAY    primal_ray_stats: a0..a6 = adx, lc, uc, dx, c, lv, uv over m + n (4)
//   RK_DRAY    dual_ray_stats: a0..a5 = dy, lc, uc, atdy, lv, uv over m + n (5)
// Workgroup b of entry op; the K sums go to dst (G = 1) or to the entry's
// partials at pbase.
enum { RK_DOT = 0, RK_DIST = 1, RK_PRIMAL = 2, RK_DUAL = 3, RK_STEP = 4, RK_PRAY = 5, RK_DRAY = 6 };

extern "C" __global__ __launch_bounds__(WG) void sp_cred_any_f64(const CBlk* __restrict__ blk,
                                                                 const COp* __restrict__ ops,
                                                                 double* __restrict__ partials) {
    const CBlk k = blk[blockIdx.x];
    const COp& o = ops[k.op];
    double* part = partials + o.pbase;
    switch (o.kind) {  // uniform across the workgroup
        case RK_DOT:
            reduce_part<1>(o.n, FDot{o.a0, o.a1}, k.b, o.G, part, o.dst);
            break;
        case RK_DIST:
            reduce_part<1>(o.n, FDist{o.a0, o.a1}, k.b, o.G, part, o.dst);
            break;
        case RK_PRIMAL:
            reduce_part<2>(o.m + o.n, FPrimal{o.a0, o.a1, o.a2, o.a3, o.a4, o.a5, o.m}, k.b, o.G, part, o.dst);
            break;
        case RK_DUAL:
            reduce_part<3>(o.m + o.n, FDual{o.a0, o.a1, o.a2, o.a3, o.a4, o.a5, o.a6, o.a7, o.m}, k.b, o.G, part,
                           o.dst);
            break;
        case RK_STEP:
            reduce_part<1>(o.n, FStep{o.a0, o.a1, o.a2, o.a3, o.a4, o.s0}, k.b, o.G, part, o.dst);
            break;
        case RK_PRAY:
            reduce_part<4>(o.m + o.n, FPrimalRay{o.a0, o.a1, o.a2, o.a3, o.a4, o.a5, o.a6, o.m}, k.b, o.G, part,
                           o.dst);
            break;
        default:
            reduce_part<5>(o.m + o.n, FDualRay{o.a0, o.a1, o.a2, o.a3, o.a4, o.a5, o.m}, k.b, o.G, part, o.dst);
            break;
    }
}

// The finishing pass of sp_cred_any_f64: one workgroup per listed entry (those
// with G > 1), its K sums from the entry.
extern "C" __global__ __launch_bounds__(WG) void sp_cred_finish_any(const int* __restrict__ which,
                                                                    const COp* __restrict__ ops,
                                                                    const double* __restrict__ partials) {
    const COp& o = ops[which[blockIdx.x]];
    red_finish_body(partials + o.pbase, o.G, o.K, o.dst);
}

// sp_chalpern_f64 with w and ρ as arguments (the same for every entry).
extern "C" __global__ __launch_bounds__(WG) void sp_chalpern_wr_f64(const CBlk* __restrict__ blk,
                                                                    const COp* __restrict__ ops, double w,
                                                                    double rho) {
    CELEM_HEAD
    for (i64 i = i0 + threadIdx.x; i < i1; i += WG)
        o.o0[i] = w * ((1.0 + rho) * o.a0[i] - rho * o.o0[i]) + (1.0 - w) * o.a1[i];
}
