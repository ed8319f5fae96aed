// Host-side model of the LP and its preconditioning, in C++.
//
// Restates crates/saddlepoint-core/src/problem.rs (CsrMatrix, LpProblem) and
// crates/saddlepoint-core/src/scaling.rs (ScaledProblem) so that the PDHG
// driver in solver.hpp can run over the HIP backend without the Rust crates.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace sp {

// Errors cross the C ABI as a code plus a message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// CsrMatrix (problem.rs:16-24) with usize indices as int64.
struct Csr {
    int64_t nrows = 0, ncols = 0;
    std::vector<int64_t> row_ptr, col_idx;
    std::vector<double> values;
    int64_t nnz() const { return (int64_t)values.size(); }
    Csr transpose() const;                                          // problem.rs:88-109
    void scale(const std::vector<double>& row, const std::vector<double>& col);  // :120-126
};

// LpProblem (problem.rs:131-145), borrowed views of the caller's arrays.
struct LpView {
    int64_t m = 0, n = 0, nnz = 0;
    const int64_t *row_ptr = nullptr, *col_idx = nullptr;
    const double* values = nullptr;
    const double *c = nullptr, *lc = nullptr, *uc = nullptr, *lv = nullptr, *uv = nullptr;
    double c0 = 0.0;
    bool maximize = false;
};

// ScalingParams (scaling.rs:15-34).
struct ScalingParams {
    int64_t geometric_mean_passes = 12;
    int64_t ruiz_iterations = 10;
    bool has_pock_chambolle = true;
    double pock_chambolle_alpha = 1.0;
    bool bound_objective_rescaling = true;
};

// ScaledProblem (scaling.rs:37-50).
struct ScaledProblem {
    Csr a, at;
    std::vector<double> c, lc, uc, lv, uv, row_scale, col_scale;
    double bound_scale = 1.0, obj_scale = 1.0;

    static ScaledProblem build(const LpView& p, const ScalingParams& params);  // :53-103
    std::vector<double> unscale_primal(const std::vector<double>& xs) const;    // :114-116
    std::vector<double> unscale_dual(const std::vector<double>& ys) const;      // :119-121
    std::vector<double> primal_residual_factors() const;                        // :124-126
    std::vector<double> dual_residual_factors() const;                          // :129-131
    double objective_factor() const { return 1.0 / (bound_scale * obj_scale); }  // :134-136
};

// The explicit transpose of a CSR matrix given by its arrays (problem.rs:88-109).
Csr transpose_of(int64_t nrows, int64_t ncols, const int64_t* row_ptr, const int64_t* col_idx, const double* values);

// The vector part of ScaledProblem::new once the matrix scaling is known
// (scaling.rs:80-103): c, bounds and the bound/objective rescaling. With
// `rank` (the position of original column j), ‖c‖ is summed in the original
// column order, so a relabeled LP scales to exactly the relabeled result.
void finish_scaling(const LpView& p, const ScalingParams& params, ScaledProblem& s, const int64_t* rank = nullptr);

// A relabeling of the columns by decreasing nonzero count (ties by index), so
// that the most gathered entries of the column-indexed vectors share cache
// lines. Each row keeps its nonzeros in their order: products and row sums
// are unchanged, only where column-indexed values are stored.
struct ColumnOrder {
    std::vector<int64_t> order;  // new position -> original column
    std::vector<int64_t> rank;   // original column -> new position
    bool empty() const { return order.empty(); }
};
ColumnOrder popularity_order(const LpView& p);
// The same order from the column counts.
ColumnOrder popularity_order_of_counts(const std::vector<int64_t>& count);
// The LP with relabeled columns; `v` points into the owned arrays.
struct RelabeledLp {
    LpView v;
    std::vector<int64_t> col_idx;
    std::vector<double> c, lv, uv;
};
RelabeledLp relabel(const LpView& p, const ColumnOrder& o);
// Only c, lv and uv relabeled (col_idx still the original's): for a caller
// that relabels the column indices on the device.
RelabeledLp relabel_vectors(const LpView& p, const ColumnOrder& o);
// The same relabeling applied to a scaled problem built from the original LP.
void relabel(ScaledProblem& s, const ColumnOrder& o);

// A relabeling of the rows, for LPs run with column panels and for large LPs
// (sp_params.reorder_rows): the rows with
// more than `long_row` nonzeros first, longest first (ties by index), then
// the others by their smallest column index (relabeled by `cols` when it is
// not empty; an empty row last), ties by index. The long rows' entries of the
// row-indexed vectors, which the A'ᵀ side gathers for most of its nonzeros
// when the long rows hold most of them, then share a few cache lines (the
// most gathered first, so that the first H of them can be kept in LDS: see
// Hip::plan_hot), and the short rows form contiguous stream blocks whose
// gathers sweep x̄ in order. Each row
// keeps its nonzeros and each column of A'ᵀ its order, so every row sum and
// every A'ᵀ product is unchanged; only sums over rows change their order.
struct RowOrder {
    std::vector<int64_t> order;  // new position -> original row
    bool empty() const { return order.empty(); }
};
RowOrder long_rows_first(const LpView& p, const ColumnOrder& cols, int64_t long_row);
// reorder_rows = -1 relabels the rows of LPs with at least this many nonzeros.
constexpr int64_t kReorderRowsMinNnz = int64_t(1) << 20;

}  // namespace sp

// An LP read from a file (the C ABI's sp_lp): LpProblem (problem.rs:131-145)
// with owned arrays.
struct sp_lp {
    std::string name;
    int64_t m = 0, n = 0;
    std::vector<int64_t> row_ptr, col_idx;
    std::vector<double> values, c, lc, uc, lv, uv;
    double c0 = 0.0;
    bool maximize = false;
};

namespace sp {

// MPS text to an LP (crates/saddlepoint-io/src/lib.rs:59-309); throws Error.
sp_lp* read_mps_text(const char* data, size_t size);

double norm2(const double* v, int64_t n);                                   // scaling.rs:186-188
double finite_bound_norm(const double* lo, const double* hi, int64_t m);    // scaling.rs:191-201
void validate(const LpView& p);                                             // problem.rs:166-194

}  // namespace sp
