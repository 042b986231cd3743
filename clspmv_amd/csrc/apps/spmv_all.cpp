// spmv_all: every kernel variant of one storage format on one matrix.
//
//     spmv_all matrix.mtx method [ntimes]
//
// Same command line and method numbering as the reference's spmv_all
// (A/src/spmv_all.cpp:158-255): 0 bandwidth and launch-overhead probes,
// 1 CSR scalar, 2 CSR vector, 3 BDIA, 4 DIA, 5 ELL, 6 COO, 7 BELL, 8 BCSR,
// 9 SELL, 10 SBELL; plus 11 = native r x c BSR (r = c = 2, 3, 4).  For each
// variant: one launch from y = 0 checked against the host FP64 product,
// then ntimes timed launches; the best variant is reported last.  x = 1 and
// y = 0 as in the reference's drivers (initVectorOne / initVectorZero).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "cli_common.hpp"

using namespace cli;

namespace {

struct Run {
    clspmv_ctx_t ctx;
    const clspmv_coo_matrix* coo;
    DevVec *x, *y;
    const Check* chk;
    int ntimes;
    double best_ms = 1e30;
    std::string best;
};

void report(Run& r, const char* fmt, const std::string& variant, clspmv_part_t part) {
    Timing t = check_and_time(r.ctx, part, *r.x, *r.y, *r.chk, r.ntimes);
    clspmv_part_info inf;
    clspmv_part_get_info(part, &inf);
    const double s = t.ms * 1e-3;
    std::printf("%-6s %-22s kernel %-22s time %9.4f ms  GFLOPS %8.2f  GB/s %8.1f  max_rel_err %.3e%s\n", fmt,
                variant.c_str(), inf.kernel, t.ms, 2.0 * inf.nnz / s / 1e9, inf.alg_bytes / s / 1e9, t.err,
                t.err <= 1e-5 ? "" : "  CHECK FAILED");
    if (t.ms < r.best_ms) {
        r.best_ms = t.ms;
        r.best = std::string(fmt) + " " + variant;
    }
    clspmv_part_free(part);
}

template <class F>
void try_part(Run& r, const char* fmt, const std::string& variant, F make) {
    clspmv_part_t part = nullptr;
    const int rc = make(&part);
    if (rc != CLSPMV_OK) {
        std::printf("%-6s %-22s not run: %s\n", fmt, variant.c_str(), clspmv_last_error());
        return;
    }
    report(r, fmt, variant, part);
}

void probes(clspmv_ctx_t ctx) {
    const size_t bytes = 256u << 20;
    void *a = nullptr, *b = nullptr;
    ok(clspmv_device_alloc(ctx, bytes, &a), "device_alloc");
    ok(clspmv_device_alloc(ctx, bytes, &b), "device_alloc");
    float gbps = 0.f, us = 0.f;
    ok(clspmv_probe_copy(ctx, a, b, bytes, 20, &gbps), "probe_copy");
    ok(clspmv_probe_launch(ctx, 1000, &us), "probe_launch");
    std::printf("device copy (256 MiB read + write): %.1f GB/s\nempty-kernel launch: %.2f us per launch\n", gbps, us);
    clspmv_device_free(ctx, a);
    clspmv_device_free(ctx, b);
}

const int kShapes[8][2] = {{1, 4}, {2, 4}, {4, 4}, {8, 4}, {1, 8}, {2, 8}, {4, 8}, {8, 8}};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::printf(
            "Usage: spmv_all input_matrix.mtx method execution_times\n"
            "  method 0 probes; 1 CSR scalar; 2 CSR vector; 3 BDIA; 4 DIA; 5 ELL; 6 COO;\n"
            "         7 BELL; 8 BCSR; 9 SELL; 10 SBELL; 11 native r x c BSR\n");
        return 0;
    }
    const int method = argc > 2 ? std::atoi(argv[2]) : 1;
    const int ntimes = argc > 3 ? std::atoi(argv[3]) : 20;
    if (method < 0 || method > 11 || ntimes < 1) {
        std::fprintf(stderr, "bad method or execution_times\n");
        return 1;
    }
    clspmv_coo_matrix coo;
    clspmv_init_coo(&coo);
    ok(clspmv_read_mmf(argv[1], &coo), "read_mmf");
    std::printf("matrix %s: width %d height %d nnz %d\n", argv[1], coo.matinfo.width, coo.matinfo.height,
                coo.matinfo.nnz);
    clspmv_ctx_t ctx;
    ok(clspmv_init(0, &ctx), "init");
    if (method == 0) {
        probes(ctx);
        clspmv_finalize(ctx);
        clspmv_free_coo(&coo);
        return 0;
    }
    std::vector<float> xh((size_t)coo.matinfo.width, 1.0f);
    Check chk(coo, xh);
    DevVec x(ctx, (size_t)coo.matinfo.width), y(ctx, (size_t)coo.matinfo.height);
    x.set(xh);
    Run r{ctx, &coo, &x, &y, &chk, ntimes};

    switch (method) {
        case 1:
        case 2: {
            clspmv_csr_matrix csr;
            ok(clspmv_coo2csr(&coo, &csr), "coo2csr");
            if (method == 1) {
                try_part(r, "CSR", "scalar", [&](clspmv_part_t* p) { return clspmv_part_csr_scalar(ctx, &csr, 0, p); });
            } else {
                const char* names[] = {"vector L lanes", "adaptive", "column-sliced", "", "LDS x-tiled"};
                for (int m : {0, 1, 2, 4})
                    try_part(r, "CSR", names[m], [&](clspmv_part_t* p) { return clspmv_part_csr_vector(ctx, &csr, m, p); });
            }
            clspmv_free_csr(&csr);
            break;
        }
        case 3: {
            clspmv_bdia_matrix b;
            if (clspmv_coo2bdia(&coo, &b, CLSPMV_GPU_ALIGNMENT) != CLSPMV_OK) {
                std::printf("BDIA: conversion refused (%s)\n", clspmv_last_error());
                break;
            }
            for (int m : {0, 1})
                try_part(r, "BDIA", m ? "LDS x band" : "direct x", [&](clspmv_part_t* p) { return clspmv_part_bdia(ctx, &b, m, p); });
            clspmv_free_bdia(&b);
            break;
        }
        case 4: {
            clspmv_dia_matrix d;
            if (clspmv_coo2dia(&coo, &d, CLSPMV_GPU_ALIGNMENT) != CLSPMV_OK) {
                std::printf("DIA: conversion refused (%s)\n", clspmv_last_error());
                break;
            }
            for (int m : {0, 1})
                try_part(r, "DIA", m ? "4 rows per lane" : "1 row per lane", [&](clspmv_part_t* p) { return clspmv_part_dia(ctx, &d, m, p); });
            clspmv_free_dia(&d);
            break;
        }
        case 5: {
            clspmv_ell_matrix e;
            ok(clspmv_coo2ell(&coo, &e, CLSPMV_GPU_ALIGNMENT, 0), "coo2ell");
            const char* names[] = {"1 row per lane", "4 rows per lane", "column-sliced", "", "LDS x-tiled"};
            for (int m : {0, 1, 2, 4})
                try_part(r, "ELL", names[m], [&](clspmv_part_t* p) { return clspmv_part_ell(ctx, &e, m, p); });
            clspmv_free_ell(&e);
            break;
        }
        case 6:
            for (int m : {0, 4})
                try_part(r, "COO", m ? "LDS x-tiled" : "segmented", [&](clspmv_part_t* p) { return clspmv_part_coo(ctx, &coo, m, p); });
            break;
        case 7:
        case 8:
        case 10:
            for (auto& sh : kShapes) {
                const int bh = sh[0], bw = sh[1];
                const std::string v = std::to_string(bh) + "x" + std::to_string(bw);
                if (method == 7) {
                    clspmv_b4ell_matrix b;
                    if (clspmv_coo2b4ell(&coo, &b, bw, bh, CLSPMV_GPU_ALIGNMENT, 0) != CLSPMV_OK) {
                        std::printf("BELL   %-22s conversion refused (%s)\n", v.c_str(), clspmv_last_error());
                        continue;
                    }
                    try_part(r, "BELL", v, [&](clspmv_part_t* p) { return clspmv_part_bell(ctx, &b, 0, p); });
                    clspmv_free_b4ell(&b);
                } else if (method == 8) {
                    clspmv_b4csr_matrix b;
                    if (clspmv_coo2b4csr(&coo, &b, bw, bh, CLSPMV_GPU_ALIGNMENT) != CLSPMV_OK) {
                        std::printf("BCSR   %-22s conversion refused (%s)\n", v.c_str(), clspmv_last_error());
                        continue;
                    }
                    try_part(r, "BCSR", v, [&](clspmv_part_t* p) { return clspmv_part_bcsr(ctx, &b, 0, p); });
                    clspmv_free_b4csr(&b);
                } else {
                    clspmv_sbell_matrix b;
                    if (clspmv_coo2sbell(&coo, &b, bw, bh, CLSPMV_WARPSIZE) != CLSPMV_OK) {
                        std::printf("SBELL  %-22s conversion refused (%s)\n", v.c_str(), clspmv_last_error());
                        continue;
                    }
                    try_part(r, "SBELL", v, [&](clspmv_part_t* p) { return clspmv_part_sbell(ctx, &b, 0, p); });
                    clspmv_free_sbell(&b);
                }
            }
            break;
        case 9:
            for (int S : {32, 128}) {
                clspmv_sell_matrix s;
                ok(clspmv_coo2sell(&coo, &s, (unsigned)S), "coo2sell");
                const std::string v = "slice " + std::to_string(S);
                try_part(r, "SELL", v, [&](clspmv_part_t* p) { return clspmv_part_sell(ctx, &s, 0, p); });
                if (S == 32)
                    try_part(r, "SELL", v + " LDS x-tiled", [&](clspmv_part_t* p) { return clspmv_part_sell(ctx, &s, 4, p); });
                clspmv_free_sell(&s);
            }
            break;
        case 11:
            for (int b : {2, 3, 4}) {
                clspmv_bsr_matrix m;
                if (clspmv_coo2bsr(&coo, &m, b, b) != CLSPMV_OK) {
                    std::printf("BSR    %dx%d conversion refused (%s)\n", b, b, clspmv_last_error());
                    continue;
                }
                for (int meth : {0, 1}) {
                    const std::string v = std::to_string(b) + "x" + std::to_string(b) + (meth ? " lanes" : " sliced");
                    try_part(r, "BSR", v, [&](clspmv_part_t* p) { return clspmv_part_bsr(ctx, &m, meth, p); });
                }
                clspmv_free_bsr(&m);
            }
            break;
    }
    if (!r.best.empty())
        std::printf("------------------------------------------------------------------------\n"
                    "best: %s  %.4f ms  GFLOPS %.2f\n",
                    r.best.c_str(), r.best_ms, 2.0 * coo.matinfo.nnz / (r.best_ms * 1e-3) / 1e9);
    clspmv_finalize(ctx);
    clspmv_free_coo(&coo);
    return 0;
}
