// spmv_cocktail: Cocktail partition of one matrix, uploaded, checked and
// timed -- the drop-in for the reference's spmv_cocktail (A/src/eval.cpp
// :75-130, do_evaluation + evaluate, A/src/spmv_cocktail.cpp:1134-1455).
//
//     spmv_cocktail matrix.mtx [ntimes] [dia_priority_multiplier]
//                   [estimate_flat_partial 0|1] [estimate_flat_full 0|1]
//
// The performance tables are read from $CLSPMV_TABLES, else from
// ../tables/reference_amd next to this executable (the reference reads
// $CLSPMVPATH/benchmark).  x = 1, y = 0 as in eval.cpp.
#include <limits.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "clspmv_cocktail.h"
#include "cli_common.hpp"

using namespace cli;

namespace {

std::string tables_dir() {
    if (const char* e = std::getenv("CLSPMV_TABLES")) return e;
    char buf[PATH_MAX];
    const ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
    std::string exe = n > 0 ? std::string(buf, (size_t)n) : std::string(".");
    const size_t slash = exe.rfind('/');
    return (slash == std::string::npos ? std::string(".") : exe.substr(0, slash)) + "/../tables/reference_amd";
}

void print_partition(const clspmv_cocktail& m) {
    std::printf("Cocktail partition: width %d height %d nnz %d\n", m.mat_width, m.mat_height, m.mat_nnz);
    struct P {
        const char* name;
        unsigned char use;
        int meth, nnz;
    } parts[] = {{"bdia", m.ifusebdia, m.bdia_meth_num, m.bdia.matinfo.nnz},
                 {"dia", m.ifusedia, m.dia_meth_num, m.dia.matinfo.nnz},
                 {"sbell", m.ifusesbell, m.sbell_meth_num, m.sbell.matinfo.nnz},
                 {"bell", m.ifusebell, m.bell_meth_num, m.bell.matinfo.nnz},
                 {"bcsr", m.ifusebcsr, m.bcsr_meth_num, m.bcsr.matinfo.nnz},
                 {"sell", m.ifusesell, m.sell_meth_num, m.sell.matinfo.nnz},
                 {"ell", m.ifuseell, m.ell_meth_num, m.ell.matinfo.nnz},
                 {"csr", m.ifusecsr, m.csr_meth_num, m.csr.matinfo.nnz},
                 {"coo", m.ifusecoo, m.coo_meth_num, m.coo.matinfo.nnz}};
    for (auto& p : parts)
        if (p.use) std::printf("  use %-5s implementation no. %d  nnz %d\n", p.name, p.meth, p.nnz);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::printf("Usage: spmv_cocktail input_matrix.mtx [ntimes] [dia_priority_multiplier] "
                    "[estimate_flat_partial 0|1] [estimate_flat_full 0|1]\n");
        return 0;
    }
    const int ntimes = argc > 2 ? std::atoi(argv[2]) : 1000;
    const float dia_mult = argc > 3 ? (float)std::atof(argv[3]) : 1.0f;
    const int partial = argc > 4 ? std::atoi(argv[4]) : 1;
    const int full = argc > 5 ? std::atoi(argv[5]) : 0;
    if (ntimes < 1) {
        std::fprintf(stderr, "bad ntimes\n");
        return 1;
    }
    clspmv_coo_matrix coo;
    clspmv_init_coo(&coo);
    ok(clspmv_read_mmf(argv[1], &coo), "read_mmf");
    std::printf("matrix %s: width %d height %d nnz %d\n", argv[1], coo.matinfo.width, coo.matinfo.height,
                coo.matinfo.nnz);
    std::vector<float> xh((size_t)coo.matinfo.width, 1.0f), y0((size_t)coo.matinfo.height, 0.0f);
    Check chk(coo, xh);

    clspmv_bench_t tables = nullptr;
    const std::string dir = tables_dir();
    ok(clspmv_bench_load(dir.c_str(), &tables), ("bench_load " + dir).c_str());
    clspmv_cocktail mat;
    clspmv_init_cocktail(&mat);
    auto t0 = std::chrono::steady_clock::now();
    ok(clspmv_analyze_matrix(tables, &coo, &mat, dia_mult, partial, full, 0), "analyze_matrix");
    const double t_an = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    print_partition(mat);
    std::printf("analysis time %.3f s (tables %s)\n", t_an, dir.c_str());

    clspmv_ctx_t ctx;
    ok(clspmv_init(0, &ctx), "init");
    clspmv_gpumat_t g = nullptr;
    ok(clspmv_init_mat_kernels(ctx, &mat, xh.data(), y0.data(), 0, &g), "init_mat_kernels");
    long long alg = 4LL * coo.matinfo.width + 8LL * coo.matinfo.height;
    for (int i = 0; i < clspmv_gpumat_num_parts(g); ++i) {
        clspmv_part_info inf;
        clspmv_gpumat_part_info(g, i, &inf);
        alg += inf.stored_bytes;
        std::printf("  part %d: kernel %s nnz %lld stored bytes %lld\n", i, inf.kernel, inf.nnz, inf.stored_bytes);
    }
    ok(clspmv_do_spmv(g, 1), "do_spmv");
    std::vector<float> y(y0.size());
    ok(clspmv_cpy_result_to_cpu(g, y.data(), (int)y.size()), "cpy_result_to_cpu");
    const double err = chk.max_rel_err(y);
    std::printf("check: max_rel_err %.3e%s\n", err, err <= 1e-5 ? "" : "  CHECK FAILED");
    ok(clspmv_do_spmv(g, 2), "do_spmv");
    ok(clspmv_synchronize(ctx), "synchronize");
    ok(clspmv_event_elapsed_begin(ctx), "event");
    ok(clspmv_do_spmv(g, ntimes), "do_spmv");
    float ms = 0.f;
    ok(clspmv_event_elapsed_end(ctx, &ms), "event");
    const double t = ms / ntimes;
    std::printf("Cocktail time %.4f ms GFLOPS %.2f GB/s %.1f\n", t, 2.0 * coo.matinfo.nnz / (t * 1e-3) / 1e9,
                alg / (t * 1e-3) / 1e9);
    clspmv_free_mat_kernels(g);
    clspmv_free_cocktail(&mat);
    clspmv_bench_free(tables);
    clspmv_finalize(ctx);
    clspmv_free_coo(&coo);
    return err <= 1e-5 ? 0 : 2;
}
