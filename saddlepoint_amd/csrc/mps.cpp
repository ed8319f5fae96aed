// Free-format MPS reader: restates crates/saddlepoint-io/src/lib.rs:59-309 and
// CsrMatrix::from_triplets (crates/saddlepoint-core/src/problem.rs:33-77) for
// files at MIPLIB scale (one pass over the bytes, hashed names, counting sort).
//
// Sections: NAME, OBJSENSE (inline or on the next line), ROWS, COLUMNS
// (integer markers read and ignored: the LP relaxation), RHS (an RHS on the
// objective row sets the constant to its negation), RANGES, BOUNDS (UP, LO,
// FX, FR, MI, PL, BV, LI, UI), ENDATA. Only the first RHS, RANGES and BOUNDS
// set is used. Extra N rows are ignored.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <string_view>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "host.hpp"
#include "saddlepoint_hip.h"

namespace sp {

namespace {

struct MpsError : Error {
    MpsError(int64_t line, const std::string& m) : Error(SP_ERR_PROBLEM, "line " + std::to_string(line) + ": " + m) {}
};

std::string upper(std::string_view s) {
    std::string u(s);
    for (char& ch : u) ch = (char)std::toupper((unsigned char)ch);
    return u;
}

// 1 = maximize, 0 = minimize, -1 = not a sense word.
int sense_of(std::string_view s) {
    const std::string u = upper(s);
    if (u == "MAX" || u == "MAXIMIZE" || u == "MAXIMISE") return 1;
    if (u == "MIN" || u == "MINIMIZE" || u == "MINIMISE") return 0;
    return -1;
}

bool number(std::string_view s, double& v) {
    char buf[64];
    if (s.empty() || s.size() >= sizeof buf) return false;
    std::memcpy(buf, s.data(), s.size());
    buf[s.size()] = 0;
    char* end = nullptr;
    errno = 0;
    v = std::strtod(buf, &end);
    return end == buf + s.size() && !std::isnan(v);
}

std::string_view strip_quotes(std::string_view s) {
    while (!s.empty() && s.front() == '\'') s.remove_prefix(1);
    while (!s.empty() && s.back() == '\'') s.remove_suffix(1);
    return s;
}

}  // namespace

// problem.rs:33-77: CSR with the columns of a row sorted, duplicates summed in
// triplet order and explicit zeros dropped.
void csr_from_triplets(int64_t m, const std::vector<int64_t>& tr, const std::vector<int64_t>& tc,
                       const std::vector<double>& tv, std::vector<int64_t>& row_ptr, std::vector<int64_t>& col_idx,
                       std::vector<double>& values) {
    const size_t t = tr.size();
    std::vector<int64_t> start((size_t)m + 1, 0);
    for (int64_t r : tr) start[(size_t)r + 1] += 1;
    for (int64_t i = 0; i < m; ++i) start[(size_t)i + 1] += start[(size_t)i];
    std::vector<int64_t> by_row(t);
    {
        std::vector<int64_t> next(start.begin(), start.end() - 1);
        for (size_t k = 0; k < t; ++k) by_row[(size_t)next[(size_t)tr[k]]++] = (int64_t)k;
    }
    row_ptr.assign((size_t)m + 1, 0);
    col_idx.clear();
    values.clear();
    col_idx.reserve(t);
    values.reserve(t);
    for (int64_t i = 0; i < m; ++i) {
        auto b = by_row.begin() + start[(size_t)i], e = by_row.begin() + start[(size_t)i + 1];
        // Stable: duplicates keep their triplet order and are summed left to right.
        std::stable_sort(b, e, [&](int64_t x, int64_t y) { return tc[(size_t)x] < tc[(size_t)y]; });
        for (auto it = b; it != e;) {
            const int64_t col = tc[(size_t)*it];
            double s = tv[(size_t)*it];
            for (++it; it != e && tc[(size_t)*it] == col; ++it) s += tv[(size_t)*it];
            if (s != 0.0) {
                col_idx.push_back(col);
                values.push_back(s);
            }
        }
        row_ptr[(size_t)i + 1] = (int64_t)col_idx.size();
    }
}

sp_lp* read_mps_text(const char* data, size_t size) {
    auto lp = std::make_unique<sp_lp>();
    enum Section { kNone, kName, kObjSense, kRows, kColumns, kRhs, kRanges, kBounds };
    Section section = kNone;
    // Names are views into `data`, which outlives the parse.
    std::string_view obj_row;
    bool have_obj = false;
    std::unordered_map<std::string_view, int64_t> row_index, col_index;
    std::unordered_set<std::string_view> ignored;
    std::vector<char> row_kind;
    std::vector<double> rhs, rng;
    std::vector<char> has_rng;
    std::vector<int64_t> tr, tc;
    std::vector<double> tv;
    std::string_view rhs_set, ranges_set, bounds_set;
    bool rhs_set_known = false, ranges_set_known = false, bounds_set_known = false;
    std::vector<std::string_view> f;
    f.reserve(16);
    int64_t lineno = 0;
    size_t pos = 0;
    while (pos < size) {
        size_t eol = pos;
        while (eol < size && data[eol] != '\n') ++eol;
        std::string_view line(data + pos, eol - pos);
        pos = eol + 1;
        ++lineno;
        if (!line.empty() && line.back() == '\r') line.remove_suffix(1);
        if (line.empty() || line[0] == '*') continue;
        f.clear();
        for (size_t i = 0; i < line.size();) {
            while (i < line.size() && (line[i] == ' ' || line[i] == '\t')) ++i;
            size_t j = i;
            while (j < line.size() && line[j] != ' ' && line[j] != '\t') ++j;
            if (j > i) f.emplace_back(line.substr(i, j - i));
            i = j;
        }
        if (f.empty()) continue;
        if (line[0] != ' ' && line[0] != '\t') {
            const std::string head = upper(f[0]);
            if (head == "NAME") {
                lp->name = f.size() > 1 ? std::string(f[1]) : std::string();
                section = kName;
            } else if (head == "OBJSENSE") {
                if (f.size() > 1) {
                    const int s = sense_of(f[1]);
                    if (s < 0) throw MpsError(lineno, "bad OBJSENSE " + std::string(f[1]));
                    lp->maximize = s == 1;
                }
                section = kObjSense;
            } else if (head == "OBJSENCE") {
                section = kObjSense;
            } else if (head == "ROWS") {
                section = kRows;
            } else if (head == "COLUMNS") {
                section = kColumns;
            } else if (head == "RHS") {
                section = kRhs;
            } else if (head == "RANGES") {
                section = kRanges;
            } else if (head == "BOUNDS") {
                section = kBounds;
            } else if (head == "ENDATA") {
                break;
            } else if ((head == "OBJSENSE_MAX" || head == "MAX" || head == "MAXIMIZE" || head == "MIN" ||
                        head == "MINIMIZE") &&
                       section == kObjSense) {
                lp->maximize = sense_of(head) == 1;
            } else if (head == "QUADOBJ" || head == "QMATRIX" || head == "QSECTION" || head == "CSECTION" ||
                       head == "SOS" || head == "INDICATORS") {
                throw MpsError(lineno, "section " + std::string(f[0]) + " is not supported");
            } else {
                throw MpsError(lineno, "unknown section " + std::string(f[0]));
            }
            continue;
        }
        switch (section) {
            case kObjSense: {
                const int s = sense_of(f[0]);
                if (s < 0) throw MpsError(lineno, "bad OBJSENSE " + std::string(f[0]));
                lp->maximize = s == 1;
                break;
            }
            case kRows: {
                if (f.size() < 2) throw MpsError(lineno, "ROWS line needs a type and a name");
                const std::string t = upper(f[0]);
                const std::string_view nm = f[1];
                if (t == "N") {
                    if (!have_obj) {
                        obj_row = nm;
                        have_obj = true;
                    } else {
                        ignored.insert(nm);
                    }
                    break;
                }
                if (t != "E" && t != "L" && t != "G") throw MpsError(lineno, "unknown row type " + t);
                if (!row_index.emplace(nm, (int64_t)row_kind.size()).second)
                    throw MpsError(lineno, "duplicate row " + std::string(nm));
                row_kind.push_back(t[0]);
                rhs.push_back(0.0);
                rng.push_back(0.0);
                has_rng.push_back(0);
                break;
            }
            case kColumns: {
                if (f.size() >= 3 && upper(strip_quotes(f[1])) == "MARKER") {
                    const std::string mk = upper(strip_quotes(f[2]));
                    if (mk != "INTORG" && mk != "INTEND") throw MpsError(lineno, "unknown marker " + mk);
                    break;
                }
                if (f.size() < 3 || f.size() % 2 == 0)
                    throw MpsError(lineno, "COLUMNS line needs a column and row/value pairs");
                const std::string_view col = f[0];
                auto it = col_index.find(col);
                int64_t j;
                if (it == col_index.end()) {
                    j = (int64_t)lp->c.size();
                    col_index.emplace(col, j);
                    lp->c.push_back(0.0);
                    lp->lv.push_back(0.0);
                    lp->uv.push_back(INFINITY);
                } else {
                    j = it->second;
                }
                for (size_t k = 1; k + 1 < f.size(); k += 2) {
                    double v;
                    if (!number(f[k + 1], v)) throw MpsError(lineno, "bad number " + std::string(f[k + 1]));
                    const std::string_view row = f[k];
                    if (have_obj && row == obj_row) {
                        lp->c[(size_t)j] += v;
                    } else {
                        auto r = row_index.find(row);
                        if (r != row_index.end()) {
                            tr.push_back(r->second);
                            tc.push_back(j);
                            tv.push_back(v);
                        } else if (!ignored.count(row)) {
                            throw MpsError(lineno, "unknown row " + std::string(row));
                        }
                    }
                }
                break;
            }
            case kRhs:
            case kRanges: {
                std::string_view set;
                size_t first = 0;
                if (f.size() % 2 == 1) {
                    set = f[0];
                    first = 1;
                }
                std::string_view& known_set = section == kRhs ? rhs_set : ranges_set;
                bool& known = section == kRhs ? rhs_set_known : ranges_set_known;
                if (!known) {
                    known_set = set;
                    known = true;
                }
                if (known_set != set) break;
                for (size_t k = first; k + 1 < f.size(); k += 2) {
                    double v;
                    if (!number(f[k + 1], v)) throw MpsError(lineno, "bad number " + std::string(f[k + 1]));
                    const std::string_view row = f[k];
                    if (have_obj && row == obj_row) {
                        if (section == kRhs) lp->c0 = -v;
                    } else {
                        auto r = row_index.find(row);
                        if (r != row_index.end()) {
                            if (section == kRhs) {
                                rhs[(size_t)r->second] = v;
                            } else {
                                rng[(size_t)r->second] = v;
                                has_rng[(size_t)r->second] = 1;
                            }
                        } else if (!ignored.count(row)) {
                            throw MpsError(lineno, "unknown row " + std::string(row));
                        }
                    }
                }
                break;
            }
            case kBounds: {
                const std::string kind = upper(f[0]);
                const bool needs_value = !(kind == "FR" || kind == "MI" || kind == "PL" || kind == "BV");
                const size_t nf = f.size();
                std::string_view set, col, val;
                bool has_val = false;
                if ((nf == 4 && needs_value) || (nf == 3 && !needs_value)) {
                    set = f[1];
                    col = f[2];
                    if (nf > 3) val = f[3], has_val = true;
                } else if ((nf == 3 && needs_value) || (nf == 2 && !needs_value)) {
                    col = f[1];
                    if (nf > 2) val = f[2], has_val = true;
                } else if (nf == 4 && !needs_value) {
                    set = f[1];
                    col = f[2];
                } else {
                    throw MpsError(lineno, "malformed BOUNDS line");
                }
                if (!bounds_set_known) {
                    bounds_set = set;
                    bounds_set_known = true;
                }
                if (bounds_set != set) break;
                auto it = col_index.find(col);
                if (it == col_index.end()) throw MpsError(lineno, "unknown column " + std::string(col));
                const size_t j = (size_t)it->second;
                double v = 0.0;
                if (has_val && !number(val, v)) throw MpsError(lineno, "bad number " + std::string(val));
                if (kind == "UP" || kind == "UI") {
                    lp->uv[j] = v;
                    // A negative upper bound with the default lower bound makes
                    // the lower bound -inf (io/src/lib.rs:241-245).
                    if (v < 0.0 && lp->lv[j] == 0.0) lp->lv[j] = -INFINITY;
                } else if (kind == "LO" || kind == "LI") {
                    lp->lv[j] = v;
                } else if (kind == "FX") {
                    lp->lv[j] = v;
                    lp->uv[j] = v;
                } else if (kind == "FR") {
                    lp->lv[j] = -INFINITY;
                    lp->uv[j] = INFINITY;
                } else if (kind == "MI") {
                    lp->lv[j] = -INFINITY;
                } else if (kind == "PL") {
                    lp->uv[j] = INFINITY;
                } else if (kind == "BV") {
                    lp->lv[j] = 0.0;
                    lp->uv[j] = 1.0;
                } else {
                    throw MpsError(lineno, "unsupported bound type " + kind);
                }
                break;
            }
            case kName:
                break;
            case kNone:
                throw MpsError(lineno, "data before the first section");
        }
    }
    // Row bounds from RHS and RANGES (io/src/lib.rs:276-289).
    const int64_t m = (int64_t)row_kind.size();
    lp->lc.resize((size_t)m);
    lp->uc.resize((size_t)m);
    for (int64_t i = 0; i < m; ++i) {
        const double b = rhs[(size_t)i], r = rng[(size_t)i];
        const char k = row_kind[(size_t)i];
        double lo, hi;
        if (!has_rng[(size_t)i]) {
            lo = k == 'L' ? -INFINITY : b;
            hi = k == 'G' ? INFINITY : b;
        } else if (k == 'E') {
            lo = r >= 0.0 ? b : b + r;
            hi = r >= 0.0 ? b + r : b;
        } else if (k == 'L') {
            lo = b - std::fabs(r);
            hi = b;
        } else {
            lo = b;
            hi = b + std::fabs(r);
        }
        lp->lc[(size_t)i] = lo;
        lp->uc[(size_t)i] = hi;
    }
    lp->m = m;
    lp->n = (int64_t)lp->c.size();
    if (lp->maximize) {
        for (double& v : lp->c) v = -v;
        lp->c0 = -lp->c0;
    }
    csr_from_triplets(m, tr, tc, tv, lp->row_ptr, lp->col_idx, lp->values);
    return lp.release();
}

}  // namespace sp
