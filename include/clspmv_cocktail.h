/*
 * clspmv_cocktail.h -- the Cocktail analyzer and the Cocktail GPU driver
 * (C ABI of libclspmv.so, continued from clspmv.h).
 *
 * Replaces (paths relative to the reference tree, A/ = clSpMV_AMD/):
 *   A/include/analyze.h:12-257      benchmark tables, cocktail<> struct,
 *                                   analyze_matrix
 *   A/src/analyze.cpp:16-210        ReadBench (reads $CLSPMVPATH/benchmark)
 *   A/include/spmv_cocktail.h:7-175 cocktail_gpu / cocktail_kernels and the
 *                                   init_mat_kernels / do_spmv / cpy_* API
 */
#ifndef CLSPMV_COCKTAIL_H_
#define CLSPMV_COCKTAIL_H_

#include "clspmv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* cocktail<int,int,float> (A/include/analyze.h:105-144), same layout: nine
 * one-byte flags, method numbers, group counts, shape, then the parts. */
typedef struct {
    unsigned char ifusebdia, ifusedia, ifusesbell, ifusebell, ifusebcsr, ifusesell, ifuseell, ifusecsr, ifusecoo;
    int bdia_meth_num, dia_meth_num, sbell_meth_num, bell_meth_num, bcsr_meth_num;
    int sell_meth_num, ell_meth_num, csr_meth_num, coo_meth_num;
    int csr_group_num, coo_group_num;
    int mat_width, mat_height, mat_nnz;
    clspmv_bdia_matrix bdia;
    clspmv_dia_matrix dia;
    clspmv_sbell_matrix sbell;
    clspmv_b4ell_matrix bell;
    clspmv_b4csr_matrix bcsr;
    clspmv_sell_matrix sell;
    clspmv_ell_matrix ell;
    clspmv_csr_matrix csr;
    clspmv_coo_matrix coo;
} clspmv_cocktail;

/* init_cocktail / free_cocktail (analyze.h:146-206). */
void clspmv_init_cocktail(clspmv_cocktail* mat);
void clspmv_free_cocktail(clspmv_cocktail* mat);

/* Performance tables (benchmark_all, analyze.h:91-103), read from the ten
 * .ben files of a directory in the reference's text schema
 * (ReadBench, analyze.cpp:177-210).  dir = "<CLSPMVPATH>/benchmark". */
typedef struct clspmv_bench_s* clspmv_bench_t;
int clspmv_bench_load(const char* dir, clspmv_bench_t* bench);
int clspmv_bench_free(clspmv_bench_t bench);
/* Number of entries read per table, in the order overhead, bdia, dia,
 * sbell, bell, bcsr, sell, ell, csr, coo. */
int clspmv_bench_sizes(clspmv_bench_t bench, int counts[10]);

/* maxflop_info (analyze.h:208-243): the interpolated best GFLOPS and method
 * per format for a (rows, nnz/row) point; exposed for parity tests. */
typedef struct {
    int bdia_meth_num;
    double bdia_max_flop;
    int dia_meth_num;
    double dia_max_flop;
    int sbell_meth_num[8];
    double sbell_max_flop[8];
    int bell_meth_num[8];
    double bell_max_flop[8];
    int bcsr_meth_num[8];
    double bcsr_max_flop[8];
    int sell_meth_num;
    double sell_max_flop;
    int ell_meth_num;
    double ell_max_flop;
    int csr_meth_num;
    int csr_groupnum;
    double csr_max_flop;
    int coo_meth_num;
    int coo_groupnum;
    double coo_max_flop;
    double max_dia_base_flop;
    double max_block_base_flop;
    double max_flat_base_flop;
} clspmv_maxflop_info;

/* update_info (analyze.cpp:1358): all maxflop_* for rows x nnz/row, with
 * nnz/row = (int)((float)nnz/(float)rows + 0.5). */
int clspmv_maxflop(clspmv_bench_t bench, int rows, int nnz, clspmv_maxflop_info* info);
/* find_overhead (analyze.cpp:413-428), microseconds. */
double clspmv_find_overhead(clspmv_bench_t bench, int groupnum, int threadnum);

/* analyze_matrix (analyze.h:257, analyze.cpp:3714-3900): splits `coo` into
 * up to nine parts; an unsorted `coo` is sorted in place first.
 * Verbose decision logging goes to stderr when verbose != 0. */
int clspmv_analyze_matrix(clspmv_bench_t bench, clspmv_coo_matrix* coo, clspmv_cocktail* mat,
                          float dia_priority_multiplier, int estimate_flat_partial, int estimate_flat_full,
                          int verbose);

/* ------------------------------------------------------------------ */
/* Cocktail GPU driver: cocktail_gpu + cocktail_kernels
 * (A/include/spmv_cocktail.h:7-161).  One object owns the device copies of
 * every active part, x (padded to a multiple of 8 floats, as cpy_shared
 * does, A/src/spmv_cocktail.cpp:230-250), y (padded to a multiple of 1024)
 * and each part's kernel choice. */
typedef struct clspmv_gpumat_s* clspmv_gpumat_t;

/* init_mat_kernels (spmv_cocktail.cpp:1590-1611): uploads the parts, x and
 * y; use_tex is accepted and ignored (no texture path on CDNA4). */
int clspmv_init_mat_kernels(clspmv_ctx_t ctx, const clspmv_cocktail* mat, const float* vec, const float* res,
                            int use_tex, clspmv_gpumat_t* gpumat);
/* free_mat_kernels (:1613-1618). */
int clspmv_free_mat_kernels(clspmv_gpumat_t gpumat);
/* do_spmv (:1088-1132): ntimes * (y += A x), parts in the reference's fixed
 * order BDIA, DIA, SBELL, BELL, BCSR, SELL, ELL, CSR, COO on one stream;
 * a plan made only of row runs (each part writing rows of its own) runs
 * its parts side by side on up to four streams, forked from and joined
 * back to the context stream ("cocktail_streams" 0: one stream);
 * asynchronous (clspmv_synchronize waits).  With
 * clspmv_set_option("cocktail_graph", 1) the passes are replayed from
 * hipGraphs captured on first use (up to 64 passes per graph launch);
 * by default each part is launched directly (measured no slower). */
int clspmv_do_spmv(clspmv_gpumat_t gpumat, int ntimes);
/* The iteration of the reference's example (A/src/eval.cpp:19-72, with
 * cpy_vector_from_gpu :1510-1556): `iterations` times { y += A x; x := y },
 * all enqueued on the device with no host round trip (x := y by a copy
 * kernel; from captured graphs of up to 64 iterations with
 * "cocktail_graph" 1).  Square matrices only; asynchronous. */
int clspmv_power_iterate(clspmv_gpumat_t gpumat, int iterations);
/* cpy_vector_from_cpu / _from_gpu, cpy_result_from_cpu / _from_gpu /
 * _to_cpu / _to_gpu (:1458-1588).  *_gpu take device pointers. */
int clspmv_cpy_vector_from_cpu(clspmv_gpumat_t gpumat, const float* vec, int vec_size);
int clspmv_cpy_vector_from_gpu(clspmv_gpumat_t gpumat, const void* d_vec, int vec_size);
int clspmv_cpy_result_from_cpu(clspmv_gpumat_t gpumat, const float* res, int res_size);
int clspmv_cpy_result_from_gpu(clspmv_gpumat_t gpumat, const void* d_res, int res_size);
int clspmv_cpy_result_to_cpu(clspmv_gpumat_t gpumat, float* res, int res_size);
int clspmv_cpy_result_to_gpu(clspmv_gpumat_t gpumat, void* d_res, int res_size);
/* Device pointers of the object's x and y (the reference reads them as
 * gpumat.vec / gpumat.res, e.g. A/src/eval.cpp:50-51). */
int clspmv_gpumat_buffers(clspmv_gpumat_t gpumat, void** d_vec, void** d_res);
/* ------------------------------------------------------------------ */
/* MI355X mode of the analyzer: the plan.  A Cocktail partition keeps the
 * reference's nine parts; on MI355X the flat ones (ELL, SELL, CSR, COO) are
 * merged into one set of rows and cut into row runs by class, classified
 * on blocks of 1024-32768 rows with boundaries placed to 128 rows: rows on
 * <= 64 well-filled diagonals as DIA, rows of dense aligned 4x4 (2x2)
 * blocks as BSR, near-equal rows as SELL, other rows as CSR (adaptive row
 * blocks, or the LDS x-tiled / column-sliced layout for scattered columns);
 * each run's layout follows its class and size, and a matrix under
 * "cocktail_small_plan" entries whose rows fall into several runs runs as
 * one CSR part.  The plan lists the parts that will run, each with its
 * format, kernel method and row range; clspmv_init_mat_kernels executes
 * the plan of the partition it is given.  Extends analyze_matrix
 * (A/src/analyze.cpp:3714-3900), whose partition stays bit-exact. */
typedef struct clspmv_plan_s* clspmv_plan_t;
typedef struct {
    int source;       /* 0..8: part of the partition (bdia, dia, sbell, bell, bcsr, sell, ell, csr, coo);
                         9: a row run of its merged flat rows */
    int format;       /* CLSPMV_FMT_* the part runs as */
    int method;       /* the method its clspmv_part_* call receives (CSR 4: LDS x-tiled, 2: column-sliced,
                         1: adaptive; BSR 0: sliced, 1: lanes per block row; DIA 4: lanes per row) */
    int row0, row1;   /* rows it covers */
    long long nnz;
    int block;        /* BSR row runs: the square block size of the part (its class's, or 2 for 4x4
                         blocks run with lanes per block row); 0 otherwise */
} clspmv_plan_part;
/* analyze_matrix (these tables and knobs) and the plan of its partition;
 * the plan owns the partition. */
int clspmv_analyze_plan(clspmv_bench_t bench, clspmv_coo_matrix* coo, float dia_priority_multiplier,
                        int estimate_flat_partial, int estimate_flat_full, clspmv_plan_t* plan);
/* The plan of a caller's partition (not copied: it must outlive the plan). */
int clspmv_plan_cocktail(const clspmv_cocktail* mat, clspmv_plan_t* plan);
int clspmv_plan_num_parts(clspmv_plan_t plan);
int clspmv_plan_get_part(clspmv_plan_t plan, int i, clspmv_plan_part* part);
/* The partition the plan runs (owned by the plan after clspmv_analyze_plan). */
const clspmv_cocktail* clspmv_plan_partition(clspmv_plan_t plan);
int clspmv_plan_free(clspmv_plan_t plan);
/* init_mat_kernels for a plan: uploads its parts, x and y. */
int clspmv_init_plan_kernels(clspmv_ctx_t ctx, clspmv_plan_t plan, const float* vec, const float* res,
                             clspmv_gpumat_t* gpumat);

/* Per-part description: count and, for part i, its format/method/kernel,
 * nnz and stored bytes (same fields as clspmv_part_info). */
int clspmv_gpumat_num_parts(clspmv_gpumat_t gpumat);
int clspmv_gpumat_part_info(clspmv_gpumat_t gpumat, int i, clspmv_part_info* info);

/* evaluate (A/include/spmv_cocktail.h:166, A/src/spmv_cocktail.cpp:1431-1455
 * with check_and_time :1134-1154): uploads the partition on device 0 with
 * x = vec and y = res, runs one do_spmv and compares y with coores (the
 * caller's CPU y = res + A vec, e.g. spmv_only), times 1000 do_spmv and
 * prints the reference's report (two_vec_compare fields, ms, GFLOPS). */
int clspmv_evaluate(const clspmv_cocktail* mat, const float* vec, const float* res, const float* coores);

/* The same on a caller's context with ntimes timed runs, the figures
 * returned (printed only when verbose != 0).  distance, max_diff and
 * max_diff_ratio are two_vec_compare's (A/src/util.cpp:93-131; sign-blind);
 * max_rel_err = max_i |coores_i - y_i| / max(|coores_i|, |y_i|) (signed). */
typedef struct {
    double distance, max_diff, max_diff_ratio, max_rel_err;
    int max_diff_id;
    int ntimes, nparts;
    double ms_per_spmv, gflops, alg_gbps;
} clspmv_eval_report;
int clspmv_evaluate_report(clspmv_ctx_t ctx, const clspmv_cocktail* mat, const float* vec, const float* res,
                           const float* coores, int ntimes, int verbose, clspmv_eval_report* report);

#ifdef __cplusplus
}
#endif

#endif /* CLSPMV_COCKTAIL_H_ */
