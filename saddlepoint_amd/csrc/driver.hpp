// The HIP backend as the PDHG driver (solver.hpp) sees it: the Backend
// methods of sp::Hip plus owning handles for vectors and problems.
#pragma once

#include <memory>

#include "backend.hpp"

namespace sp {

struct VecDel {
    void operator()(sp_vec* v) const { free_vec(v); }
};
struct ProbDel {
    void operator()(sp_problem* p) const { free_problem(p); }
};
struct Vec32Del {
    void operator()(sp_vec32* v) const { free_vec32(v); }
};

struct HostDel {
    void operator()(double* h) const { free_host(h); }
};
struct OpsDel {
    void operator()(DevOpList* d) const { free_ops(d); }
};

struct DriverBackend : Hip {
    using Vector = std::unique_ptr<sp_vec, VecDel>;
    using Problem = std::unique_ptr<sp_problem, ProbDel>;
    using Vector32 = std::unique_ptr<sp_vec32, Vec32Del>;
    using HostBuffer = std::unique_ptr<double, HostDel>;
    using Ops = std::unique_ptr<DevOpList, OpsDel>;
    Problem load(const ScaledProblem& s, bool autotune = false) { return Problem(Hip::load(s, autotune)); }
    Problem load_scaled(const LpView& p, const ScalingParams& sp, bool autotune, ScaledProblem& s,
                        ColumnOrder* order = nullptr) {
        return Problem(Hip::load_scaled(p, sp, autotune, s, order));
    }
    Vector zeros(int64_t n) { return Vector(Hip::zeros(n)); }
    Vector upload(const double* d, int64_t n) { return Vector(Hip::upload(d, n)); }
    Vector32 zeros32(int64_t n) { return Vector32(Hip::zeros32(n)); }
    HostBuffer host_buffer(int64_t n) { return HostBuffer(Hip::host_alloc(n)); }
    Ops upload_ops(const OpList& l) { return Ops(Hip::upload_ops(l)); }
};

}  // namespace sp
