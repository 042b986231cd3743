/*
 * clspmv_batch.h -- batched multi-right-hand-side SpMV (libclspmv.so).
 *
 * No counterpart in the reference (clSpMV runs one matrix and one vector at
 * a time).  Callers: iterative block solvers (k right-hand sides at once) and
 * services that apply many small and large matrices per step.  A batch holds
 * n CSR matrices, matrix i with k_i in {1, 4, 16} right-hand sides, on the
 * device; clspmv_batch_run computes Y_i <- A_i X_i + Y_i for every i in one
 * persistent kernel launch (plus one small kernel when some row is longer
 * than the long-row threshold), optionally replayed from a captured
 * hipGraph.  X_i is cols_i x k_i and Y_i is rows_i x k_i, both row-major
 * (x[col * k + j]), in device buffers owned by the batch.
 */
#ifndef CLSPMV_BATCH_H_
#define CLSPMV_BATCH_H_

#include "clspmv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clspmv_batch_s* clspmv_batch_t;

/* Work-item shape of one matrix (the autotuner's knobs): rows of the matrix
 * are grouped into items of at most rows_cap rows and nnz_cap entries
 * (nnz_cap <= 512); rows longer than long_threshold entries are cut into
 * 1024-entry chunks whose partial sums are added in order by a second pass. */
typedef struct {
    int rows_cap;
    int nnz_cap;
    int long_threshold;
} clspmv_spmm_tile;

/* Default tile for k right-hand sides. */
clspmv_spmm_tile clspmv_spmm_default_tile(int k);

/* Copies the n matrices to the device and plans their items.  tiles may be
 * NULL (defaults).  X_i and Y_i start at zero. */
int clspmv_batch_create(clspmv_ctx_t ctx, int n, const clspmv_csr_matrix* const* mats, const int* ks,
                        const clspmv_spmm_tile* tiles, clspmv_batch_t* out);
/* One matrix of a batch: a CSR matrix (bsr NULL) or a BSR matrix with
 * dense r x r blocks, r = 2..4 (csr NULL; k must be 4 or 16), whose blocks
 * run on the matrix cores (v_mfma_f32_4x4x1f32: each block times k
 * right-hand sides is a dense contraction).  The caller chooses per matrix,
 * e.g. from the format selector. */
typedef struct {
    const clspmv_csr_matrix* csr;
    const clspmv_bsr_matrix* bsr;
    int k;
} clspmv_batch_entry;
/* As clspmv_batch_create, for a mix of CSR and BSR matrices (tiles apply
 * to the CSR ones and may be NULL). */
int clspmv_batch_create_entries(clspmv_ctx_t ctx, int n, const clspmv_batch_entry* entries,
                                const clspmv_spmm_tile* tiles, clspmv_batch_t* out);
int clspmv_batch_free(clspmv_batch_t b);
int clspmv_batch_size(clspmv_batch_t b);
int clspmv_batch_buffers(clspmv_batch_t b, int i, void** d_x, void** d_y);
int clspmv_batch_set_x(clspmv_batch_t b, int i, const float* x);  /* cols_i * k_i floats */
int clspmv_batch_set_y(clspmv_batch_t b, int i, const float* y);  /* rows_i * k_i floats */
int clspmv_batch_get_y(clspmv_batch_t b, int i, float* y);
/* ntimes * (Y_i += A_i X_i for all i), asynchronous on the context stream. */
int clspmv_batch_run(clspmv_batch_t b, int ntimes);
/* Captures one run into a hipGraph; later clspmv_batch_run calls replay it. */
int clspmv_batch_capture(clspmv_batch_t b);
/* Totals: true nonzeros, nonzeros x right-hand sides, algorithmic bytes of
 * one run (8 B per entry + row pointers + X read + Y read and written),
 * work items, workgroups of the launch. */
int clspmv_batch_stats(clspmv_batch_t b, long long* nnz, long long* nnz_rhs, long long* alg_bytes, int* items,
                       int* grid);

#ifdef __cplusplus
}
#endif
#endif
