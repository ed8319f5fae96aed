// sp_dual_stats_noalloc: dual_stats for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc). It has no counterpart in the reference.
//
// sp_dual_stats reduces into the backend's reduction buffers, which the
// library allocates when the backend is created. This variant keeps those
// per-workgroup partials and the three totals in `scratch`, a file-scope
// device array of fixed size that each GPU holds from the library's load
// until the process ends, so a call makes no allocation of any kind. The
// kernels are sp_dual_stats's own (the same grid, the same order), so the
// result is bitwise sp_dual_stats's.
//
// `scratch` names a host-side stand-in in host code, not a device address:
// its address on a device comes from hipGetSymbolAddress, taken once per
// device with that device current (sp_backend_create does it for its
// device) and kept per device; host code passes only that address on.
// Every call on a GPU shares that GPU's array, and calls that overlap on one
// GPU race (issue them from one thread, or on one stream in order). Nothing
// in it needs to start at zero: every partial a call reads it has written.
#include <hip/hip_runtime.h>

#include <mutex>
#include <unordered_map>

#include "backend.hpp"

namespace {

// 1024 workgroups (the reductions' largest grid) × 8 sums, plus 8 totals.
constexpr int64_t kScratchDoubles = 1024 * 8 + 8;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mu;
std::unordered_map<int, double*> g_scratch;  // device ordinal -> its copy of `scratch`

}  // namespace

namespace sp {

double* scratch_on_device(int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch.find(device);
    if (it != g_scratch.end()) return it->second;
    int current = -1;
    hip_check(hipGetDevice(&current), "hipGetDevice");
    if (current != device) throw Error(SP_ERR_INVALID, "scratch: device " + std::to_string(device) + " is not current");
    void* ptr = nullptr;
    hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) throw Error(SP_ERR_HIP, "hipGetSymbolAddress(scratch): null address");
    g_scratch.emplace(device, static_cast<double*>(ptr));
    return static_cast<double*>(ptr);
}

int64_t scratch_doubles() { return kScratchDoubles; }

}  // namespace sp

extern "C" int sp_dual_stats_noalloc(sp_backend* b, const sp_problem* p, const sp_vec* y, const sp_vec* aty,
                                     sp_dual_stats_t* out) {
    if (!b) return SP_ERR_INVALID;
    try {
        if (!p || p->b != b) throw sp::Error(SP_ERR_INVALID, "dual_stats_noalloc: null problem or another backend's");
        if (!out) throw sp::Error(SP_ERR_INVALID, "dual_stats_noalloc: null output");
        if (!b->scratch) throw sp::Error(SP_ERR_UNAVAILABLE, "dual_stats_noalloc: no scratch on this device");
        *out = sp::Hip{b}.dual_stats_in(p, y, aty, b->scratch, kScratchDoubles);
        sp::set_error(b, "");
        return SP_OK;
    } catch (const sp::Error& e) {
        sp::set_error(b, e.what());
        return e.code;
    } catch (const std::exception& e) {
        sp::set_error(b, e.what());
        return SP_ERR_INVALID;
    }
}
