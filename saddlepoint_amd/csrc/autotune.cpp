// The autotuned SpMV plan rule.
//
// A matrix is placed in a cell by two features of its row lengths: the size
// class floor(log2(nnz)) and the mean-row-length class (< 8, < 32, < 128,
// >= 128 nonzeros). The table (tuning_rule.inc) holds, per cell, the plan that
// was fastest on the sweep's matrices of that cell, summed over A' and A'ᵀ
// shapes; saddlepoint_amd/autotune.py times the candidate plans on the GPU and
// writes the table. A cell the sweep did not cover takes the plan of the
// nearest size class with the same mean class, else the default plan. The
// rule depends only on the row pointers, so an LP gets the same plan (and the
// same summation order) wherever it runs.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "backend.hpp"

namespace sp {

namespace {

struct Cell {
    int size_class, mean_class;
    int block_nnz, long_row, max_rows;
};

const Cell kRule[] = {
#include "tuning_rule.inc"
};
constexpr int kRuleLen = (int)(sizeof(kRule) / sizeof(kRule[0]));

int mean_class(double mean) { return mean < 8 ? 0 : mean < 32 ? 1 : mean < 128 ? 2 : 3; }

}  // namespace

SpmvTuning autotuned_plan(const int64_t* rp, int64_t nrows) {
    SpmvTuning t;
    const int64_t nnz = nrows > 0 ? rp[nrows] - rp[0] : 0;
    if (nnz <= 0) return t;
    int sc = 0;
    while ((int64_t(1) << (sc + 1)) <= nnz) ++sc;
    const int mc = mean_class((double)nnz / (double)nrows);
    int best = -1, best_d = 1 << 30;
    for (int i = 0; i < kRuleLen; ++i) {
        if (kRule[i].block_nnz <= 0 || kRule[i].mean_class != mc) continue;
        const int d = std::abs(kRule[i].size_class - sc);
        if (d < best_d) best_d = d, best = i;
    }
    if (best < 0) return t;
    t.block_nnz = kRule[best].block_nnz;
    t.long_row = kRule[best].long_row;
    t.max_rows = kRule[best].max_rows;
    return t;
}

}  // namespace sp
