// Host-side memory and undefined-behaviour check of the C ABI's host paths,
// built with AddressSanitizer and UBSan (`make -C saddlepoint_amd/csrc
// host-check`; no GPU is touched): every MPS file of a directory read, its
// LP scaled on the host (ScaledProblem::new) and given an SpMV plan, a NaN
// bound rejected through the validation error path, and malformed MPS text
// rejected with a line-numbered error. Exits non-zero on any unexpected
// result; the sanitizers abort on a memory error or leak.
#include <dirent.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "saddlepoint_hip.h"

namespace {

int check_file(const std::string& path) {
    sp_lp* lp = nullptr;
    if (sp_lp_read_mps(path.c_str(), &lp) != SP_OK) {
        std::printf("%s: %s\n", path.c_str(), sp_last_error(nullptr));
        return 1;
    }
    sp_lp_view v{};
    sp_lp_view_of(lp, &v);
    const int64_t m = v.a.nrows, n = v.a.ncols, nnz = v.a.nnz;
    std::vector<double> av((size_t)nnz), c((size_t)n), lc((size_t)m), uc((size_t)m), lv((size_t)n), uv((size_t)n),
        rs((size_t)m), cs((size_t)n);
    double bs = 0, os = 0;
    sp_params p;
    sp_default_params(&p);
    int bad = 0;
    if (sp_scale_problem(&v, &p, av.data(), c.data(), lc.data(), uc.data(), lv.data(), uv.data(), rs.data(),
                         cs.data(), &bs, &os) != SP_OK ||
        !(bs > 0) || !(os > 0)) {
        std::printf("%s: scaling failed: %s\n", path.c_str(), sp_last_error(nullptr));
        bad = 1;
    }
    sp_spmv_tuning t{};
    sp_spmv_tuning_for(v.a.row_ptr, m, &t);
    if (t.block_nnz < 1 || t.block_nnz > 2048) bad = 1;
    if (n > 0) {
        std::vector<double> lv2(v.lv, v.lv + n);
        lv2[0] = NAN;
        sp_lp_view w = v;
        w.lv = lv2.data();
        if (sp_scale_problem(&w, &p, av.data(), c.data(), lc.data(), uc.data(), lv.data(), uv.data(), rs.data(),
                             cs.data(), &bs, &os) != SP_ERR_PROBLEM) {
            std::printf("%s: a NaN bound was not rejected\n", path.c_str());
            bad = 1;
        }
    }
    std::printf("%s: m=%lld n=%lld nnz=%lld %s\n", path.c_str(), (long long)m, (long long)n, (long long)nnz,
                bad ? "FAILED" : "ok");
    sp_lp_free(lp);
    return bad;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: host_check <mps directory> <scratch file>\n");
        return 2;
    }
    const std::string dir = argv[1], scratch = argv[2];
    std::vector<std::string> files;
    if (DIR* d = opendir(dir.c_str())) {
        while (dirent* e = readdir(d)) {
            const std::string s = e->d_name;
            if (s.size() > 4 && s.compare(s.size() - 4, 4, ".mps") == 0) files.push_back(dir + "/" + s);
        }
        closedir(d);
    }
    int bad = files.empty();
    for (const std::string& f : files) bad |= check_file(f);
    const char* malformed[] = {
        "NAME X\nROWS\n N COST\nCOLUMNS\n X1 COST\n",
        "ROWS\n L R1\nCOLUMNS\n X1 R9 1.0\nENDATA\n",
        "NAME\nROWS\n N C\n L R\nCOLUMNS\n X C 1 R 2 Q\nENDATA\n",
    };
    for (const char* text : malformed) {
        FILE* f = std::fopen(scratch.c_str(), "w");
        if (!f) return 2;
        std::fputs(text, f);
        std::fclose(f);
        sp_lp* lp = nullptr;
        const int rc = sp_lp_read_mps(scratch.c_str(), &lp);
        std::printf("malformed: rc %d (%s)\n", rc, rc == SP_OK ? "accepted" : sp_last_error(nullptr));
        if (rc != SP_ERR_PROBLEM) bad = 1;
        if (lp) sp_lp_free(lp);
    }
    std::printf(bad ? "host check FAILED\n" : "host check ok\n");
    return bad;
}
