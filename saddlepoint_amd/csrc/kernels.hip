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

// Stream blocks stage at most this many products in LDS; a longer row is split
// into chunks of SP_CHUNK nonzeros, one workgroup each.
#define SP_NNZ_F64 2048
#define SP_NNZ_F32 4096
#define SP_CHUNK 4096

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

// ---- cross-workgroup publication (agent scope: the 8 XCDs have separate L2s)

__device__ __forceinline__ void store_agent(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_agent(const double* p) {
    return __longlong_as_double(
        (long long)__hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void store_agent(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<unsigned*>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float load_agent(const float* p) {
    return __uint_as_float(
        __hip_atomic_load(reinterpret_cast<const unsigned*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
// Returns true in exactly one workgroup: the last of `count` to arrive. The
// ticket is reset by that workgroup so the next launch starts from zero.
__device__ __forceinline__ bool arrive_last(unsigned* ticket, unsigned count) {
    __threadfence();
    const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    const bool last = (t == count - 1);
    if (last) {
        __threadfence();
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return last;
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
// out_row = Σ_k val[k] * x[col[k]] over the row, then epi(row, sum).
// Stream blocks stage the products in LDS with fully coalesced loads and then
// reduce each row with T lanes (T a power of two chosen from the block's row
// count, so the order is fixed for a given matrix). T = 1 sums a row left to
// right, as the CPU backend does. Long rows are summed by a whole workgroup
// per chunk; chunks of one row are combined in chunk order by the last
// workgroup to finish.
template <int NNZ, class V, class X, class A, class Epi>
__device__ __forceinline__ void spmv_body(const Desc* __restrict__ desc, const int* __restrict__ rp,
                                          const int* __restrict__ ci, const V* __restrict__ va,
                                          const X* __restrict__ x, A* __restrict__ partials,
                                          unsigned* __restrict__ tickets, const Epi& epi) {
    __shared__ A lds[NNZ];
    const Desc d = desc[blockIdx.x];
    const int tid = threadIdx.x;
    if (d.r1 >= 0) {
        const int r0 = d.r0, nr = d.r1 - d.r0, k0 = d.k0, n = d.k1 - d.k0;
        constexpr int PER = NNZ / WG;
        int c[PER];
        V v[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int p = tid + u * WG;
            if (p < n) {
                c[u] = __builtin_nontemporal_load(ci + k0 + p);
                v[u] = __builtin_nontemporal_load(va + k0 + p);
            }
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int p = tid + u * WG;
            if (p < n) lds[p] = (A)v[u] * (A)x[c[u]];
        }
        __syncthreads();
        int T = WAVE;
        while (T > 1 && T * nr > WG) T >>= 1;
        const int g = tid / T, l = tid % T, ng = WG / T;
        for (int base = 0; base < nr; base += ng) {
            const int rr = base + g;
            const bool act = rr < nr;
            A s = 0;
            if (act) {
                const int a = rp[r0 + rr] - k0, b = rp[r0 + rr + 1] - k0;
                for (int q = a + l; q < b; q += T) s += lds[q];
            }
            for (int o = T / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, WAVE);
            if (act && l == 0) epi(r0 + rr, s);
        }
    } else {
        const int r = d.r0, pslot = -1 - d.r1, k0 = d.k0, n = d.k1 - d.k0;
        A s[1] = {0};
#pragma unroll 4
        for (int p = tid; p < n; p += WG) {
            const int cc = __builtin_nontemporal_load(ci + k0 + p);
            const V vv = __builtin_nontemporal_load(va + k0 + p);
            s[0] += (A)vv * (A)x[cc];
        }
        block_sum<1>(s);
        const int rs = rp[r], len = rp[r + 1] - rs;
        const int nch = (len + SP_CHUNK - 1) / SP_CHUNK;
        if (nch == 1) {
            if (tid == 0) epi(r, s[0]);
            return;
        }
        if (tid == 0) {
            const int pbase = pslot - (k0 - rs) / SP_CHUNK;
            store_agent(partials + pslot, s[0]);
            if (arrive_last(tickets + pbase, (unsigned)nch)) {
                A tot = 0;
                for (int q = 0; q < nch; ++q) tot += load_agent(partials + pbase + q);
                epi(r, tot);
            }
        }
    }
}

template <class A>
struct EpiStore {
    A* out;
    __device__ void operator()(int r, A s) const { out[r] = s; }
};

// spmv_t + primal_step: aty = A'ᵀy; x_t = clamp(x - τ(c - aty), lv, uv);
// x_bar = 2 x_t - x (cpu/src/lib.rs:147-160). Stores aty, x_t and x_bar.
struct EpiPrimal {
    const double *x, *c, *lv, *uv;
    double tau;
    double *aty, *x_t, *x_bar;
    __device__ void operator()(int j, double s) const {
        const double xj = x[j];
        const double t = clampv(xj - tau * (c[j] - s), lv[j], uv[j]);
        aty[j] = s;
        x_t[j] = t;
        x_bar[j] = 2.0 * t - xj;
    }
};

// As EpiPrimal, then the reflected Halpern update of x in place
// (cpu/src/lib.rs:177-187); x_t and aty are not stored.
struct EpiPrimalHalpern {
    double* x;
    const double *c, *lv, *uv, *anchor;
    double tau, w, rho;
    double* x_bar;
    __device__ void operator()(int j, double s) const {
        const double xj = x[j];
        const double t = clampv(xj - tau * (c[j] - s), lv[j], uv[j]);
        x_bar[j] = 2.0 * t - xj;
        x[j] = w * ((1.0 + rho) * t - rho * xj) + (1.0 - w) * anchor[j];
    }
};

// spmv + dual_step: ax = A'x_bar; y_t = y - σ(ax - clamp(ax - y/σ, lc, uc))
// (cpu/src/lib.rs:162-175). Stores ax and y_t.
struct EpiDual {
    const double *y, *lc, *uc;
    double sigma;
    double *ax, *y_t;
    __device__ void operator()(int i, double s) const {
        const double yi = y[i];
        const double v = s - yi / sigma;
        ax[i] = s;
        y_t[i] = yi - sigma * (s - clampv(v, lc[i], uc[i]));
    }
};

struct EpiDualHalpern {
    double* y;
    const double *lc, *uc, *anchor;
    double sigma, w, rho;
    __device__ void operator()(int i, double s) const {
        const double yi = y[i];
        const double v = s - yi / sigma;
        const double t = yi - sigma * (s - clampv(v, lc[i], uc[i]));
        y[i] = w * ((1.0 + rho) * t - rho * yi) + (1.0 - w) * anchor[i];
    }
};

#define SPMV_ARGS                                                                                    \
    const Desc *__restrict__ desc, const int *__restrict__ rp, const int *__restrict__ ci,          \
        const double *__restrict__ va, const double *__restrict__ xin, double *__restrict__ partials, \
        unsigned *__restrict__ tickets

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_f64(SPMV_ARGS, double* __restrict__ out) {
    spmv_body<SP_NNZ_F64, double, double, double>(desc, rp, ci, va, xin, partials, tickets, EpiStore<double>{out});
}

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_t_primal_f64(SPMV_ARGS, const double* __restrict__ x,
                                                                      const double* __restrict__ c,
                                                                      const double* __restrict__ lv,
                                                                      const double* __restrict__ uv, double tau,
                                                                      double* __restrict__ aty,
                                                                      double* __restrict__ x_t,
                                                                      double* __restrict__ x_bar) {
    spmv_body<SP_NNZ_F64, double, double, double>(desc, rp, ci, va, xin, partials, tickets,
                                                  EpiPrimal{x, c, lv, uv, tau, aty, x_t, x_bar});
}

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_t_primal_halpern_f64(
    SPMV_ARGS, double* __restrict__ x, const double* __restrict__ c, const double* __restrict__ lv,
    const double* __restrict__ uv, const double* __restrict__ anchor, double tau, double w, double rho,
    double* __restrict__ x_bar) {
    spmv_body<SP_NNZ_F64, double, double, double>(desc, rp, ci, va, xin, partials, tickets,
                                                  EpiPrimalHalpern{x, c, lv, uv, anchor, tau, w, rho, x_bar});
}

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_dual_f64(SPMV_ARGS, const double* __restrict__ y,
                                                                  const double* __restrict__ lc,
                                                                  const double* __restrict__ uc, double sigma,
                                                                  double* __restrict__ ax,
                                                                  double* __restrict__ y_t) {
    spmv_body<SP_NNZ_F64, double, double, double>(desc, rp, ci, va, xin, partials, tickets,
                                                  EpiDual{y, lc, uc, sigma, ax, y_t});
}

extern "C" __global__ __launch_bounds__(WG) void sp_spmv_dual_halpern_f64(
    SPMV_ARGS, double* __restrict__ y, const double* __restrict__ lc, const double* __restrict__ uc,
    const double* __restrict__ anchor, double sigma, double w, double rho) {
    spmv_body<SP_NNZ_F64, double, double, double>(desc, rp, ci, va, xin, partials, tickets,
                                                  EpiDualHalpern{y, lc, uc, anchor, sigma, w, rho});
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

extern "C" __global__ __launch_bounds__(WG) void sp_fill_f64(double* __restrict__ v, double value, i64 n) {
    GRID_LOOP(i, n) { v[i] = value; }
}

// ---- fused reductions (cpu/src/lib.rs:189-259) ---------------------------------
//
// Each kernel runs over a combined index range (rows, then columns when a
// method reduces both), accumulates K sums per thread in a grid-stride loop,
// reduces them per workgroup in a fixed order, and the last workgroup to
// finish adds the per-workgroup partials in workgroup order and writes the K
// totals to `result`. One launch and one 8·K-byte readback per method.

template <int K, class F>
__device__ __forceinline__ void reduce_body(i64 n, const F& f, double* __restrict__ partials,
                                            unsigned* __restrict__ ticket, double* __restrict__ result) {
    double s[K];
#pragma unroll
    for (int k = 0; k < K; ++k) s[k] = 0.0;
    GRID_LOOP(i, n) { f(i, s); }
    block_sum<K>(s);
    if (gridDim.x == 1) {
        if (threadIdx.x == 0)
#pragma unroll
            for (int k = 0; k < K; ++k) result[k] = s[k];
        return;
    }
    __shared__ bool last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) store_agent(partials + (i64)blockIdx.x * K + k, s[k]);
        last = arrive_last(ticket, gridDim.x);
    }
    __syncthreads();
    if (!last) return;
#pragma unroll
    for (int k = 0; k < K; ++k) s[k] = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += WG)
#pragma unroll
        for (int k = 0; k < K; ++k) s[k] += load_agent(partials + (i64)b * K + k);
    block_sum<K>(s);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) result[k] = s[k];
}

#define RED_TAIL double *__restrict__ partials, unsigned *__restrict__ ticket, double *__restrict__ result

struct FDot {
    const double *a, *b;
    __device__ void operator()(i64 i, double (&s)[1]) const { s[0] += a[i] * b[i]; }
};
extern "C" __global__ __launch_bounds__(WG) void sp_red_dot_f64(const double* __restrict__ a,
                                                                const double* __restrict__ b, i64 n, RED_TAIL) {
    reduce_body<1>(n, FDot{a, b}, partials, ticket, result);
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
    reduce_body<1>(n, FDist{a, b}, partials, ticket, result);
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
    reduce_body<2>(m + n, FPrimal{ax, lc, uc, pf, c, x, m}, partials, ticket, result);
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
    reduce_body<3>(m + n, FDual{y, lc, uc, c, aty, lv, uv, df, m}, partials, ticket, result);
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
    reduce_body<1>(n, FStep{x, x_t, aty, aty_t, df, inv_tau}, partials, ticket, result);
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
    reduce_body<4>(m + n, FPrimalRay{adx, lc, uc, dx, c, lv, uv, m}, partials, ticket, result);
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
    reduce_body<5>(m + n, FDualRay{dy, lc, uc, atdy, lv, uv, m}, partials, ticket, result);
}
