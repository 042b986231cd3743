// Shared helpers of the spmv_all / spmv_cocktail command-line drivers:
// error exits, device buffers, the one-launch result check and the timing
// loop (A/src/spmv_ell.cpp:80-125 and A/src/spmv_cocktail.cpp:1134-1154 do
// the same check-then-time sequence).
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "clspmv.h"

namespace cli {

inline void die(const char* what, int rc) {
    std::fprintf(stderr, "error: %s failed (%d): %s\n", what, rc, clspmv_last_error());
    std::exit(1);
}
inline void ok(int rc, const char* what) {
    if (rc != CLSPMV_OK) die(what, rc);
}

// Host check of a GPU result: y_ref = A x in FP64 over the COO entries and
// the per-row scale S_i = sum_j |a_ij x_j|; the reported error is
// max_i |y_i - y_ref_i| / max(|y_ref_i|, S_i).  (The reference prints its
// two_vec_compare distances, A/src/util.cpp:93-131; this is the
// condition-aware form of the same check.)
struct Check {
    std::vector<double> yref, scale;
    Check(const clspmv_coo_matrix& m, const std::vector<float>& x) : yref(m.matinfo.height, 0.0), scale(m.matinfo.height, 0.0) {
        for (int k = 0; k < m.matinfo.nnz; ++k) {
            const double p = (double)m.coo_data[k] * (double)x[(size_t)m.coo_col_id[k]];
            yref[(size_t)m.coo_row_id[k]] += p;
            scale[(size_t)m.coo_row_id[k]] += std::fabs(p);
        }
    }
    double max_rel_err(const std::vector<float>& y) const {
        double e = 0.0;
        for (size_t i = 0; i < y.size(); ++i) {
            const double d = std::fabs((double)y[i] - yref[i]);
            const double s = std::fmax(std::fmax(std::fabs(yref[i]), scale[i]), 1e-30);
            e = std::fmax(e, d / s);
        }
        return e;
    }
};

struct DevVec {
    clspmv_ctx_t ctx;
    void* p = nullptr;
    size_t n;
    DevVec(clspmv_ctx_t c, size_t n_) : ctx(c), n(n_) { ok(clspmv_device_alloc(c, n * 4 + 16, &p), "device_alloc"); }
    ~DevVec() { clspmv_device_free(ctx, p); }
    void set(const std::vector<float>& h) { ok(clspmv_memcpy_h2d(ctx, p, h.data(), n * 4), "h2d"); }
    std::vector<float> get() const {
        std::vector<float> h(n);
        ok(clspmv_memcpy_d2h(ctx, h.data(), p, n * 4), "d2h");
        return h;
    }
};

struct Timing {
    double ms = 0.0, err = 0.0;
};

// One checked launch from y = 0, then ntimes timed launches (HIP events on
// the context stream).
inline Timing check_and_time(clspmv_ctx_t ctx, clspmv_part_t part, DevVec& x, DevVec& y, const Check& chk,
                             int ntimes) {
    Timing t;
    std::vector<float> zero(y.n, 0.f);
    y.set(zero);
    ok(clspmv_part_spmv(part, (const float*)x.p, (float*)y.p, 1), "part_spmv");
    t.err = chk.max_rel_err(y.get());
    ok(clspmv_part_spmv(part, (const float*)x.p, (float*)y.p, 2), "part_spmv");  // warm-up
    ok(clspmv_synchronize(ctx), "synchronize");
    ok(clspmv_event_elapsed_begin(ctx), "event");
    ok(clspmv_part_spmv(part, (const float*)x.p, (float*)y.p, ntimes), "part_spmv");
    float ms = 0.f;
    ok(clspmv_event_elapsed_end(ctx, &ms), "event");
    t.ms = ms / ntimes;
    return t;
}

}  // namespace cli
