/*
 * spmv_oracle.c -- CPU oracle for the clspmv_amd SpMV engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker; the product (clspmv_amd, libclspmv.so) never calls it.
 *
 * A plain-C restatement of the reference's CPU SpMV path:
 *   oracle_spmv_only      spmv_only -> coo_spmv (A/src/util.cpp:85-91,
 *                         A/include/spmv_serial.h:184-196): y = 0, then
 *                         y[row_k] += val_k * x[col_k] in COO order, FP32,
 *                         separate multiply and add (built with
 *                         -ffp-contract=off, as the reference's x86-64 g++
 *                         build has no FMA).
 *   oracle_coo_spmv_f64   the same loop accumulated in double (FP64 oracle).
 *   oracle_abs_row_sum    S_i = sum_j |a_ij x_j| in double, the scale of the
 *                         condition-aware tolerance.
 *   oracle_<fmt>_spmv     per-format CPU SpMVs (A/include/spmv_serial.h
 *                         :12-279) and, for the GPU-only layouts b4csr/b4ell
 *                         that have no reference CPU SpMV, the layout
 *                         formulas of matrix_storage.h:1872-1895, :2150-2196;
 *                         used to check each layout's indexing.
 *   oracle_two_vec_compare  the reference's report (A/src/util.cpp:93-131).
 *   oracle_csr_spmv_omp   y = A x over CSR rows with OpenMP on all host
 *                         cores: a project CPU baseline (not the reference),
 *                         timed beside the reference's one-core SpMV.
 *
 * How each function is pinned (tests/test_io_oracle_abi.py): closed forms (5-point
 * Laplacian with x = 1 gives 0 / 1 / 2 exactly; a stencil with unit values
 * and x = 1 gives the neighbour count), brute-force dense products on small
 * matrices, agreement with the reference's own compiled spmv_only
 * (oracle/_ref) on seeded matrices, and the golden vectors in tests/golden.
 */
#include <math.h>
#include <omp.h>
#include <stddef.h>
#include <string.h>

/* spmv_only + coo_spmv: y = A*x, FP32, sequential in COO order. */
void oracle_spmv_only(int nnz, int height, const int* row, const int* col, const float* val, const float* x,
                      float* y) {
    for (int i = 0; i < height; ++i) y[i] = 0.0f;
    for (int k = 0; k < nnz; ++k) {
        float prod = val[k] * x[col[k]];
        y[row[k]] = y[row[k]] + prod;
    }
}

/* coo_spmv only (y += A*x), FP32. */
void oracle_coo_spmv(int nnz, const int* row, const int* col, const float* val, const float* x, float* y) {
    for (int k = 0; k < nnz; ++k) {
        float prod = val[k] * x[col[k]];
        y[row[k]] = y[row[k]] + prod;
    }
}

/* FP64 oracle: y64 = A*x with products and sums in double. */
void oracle_coo_spmv_f64(int nnz, int height, const int* row, const int* col, const float* val, const float* x,
                         double* y) {
    for (int i = 0; i < height; ++i) y[i] = 0.0;
    for (int k = 0; k < nnz; ++k) y[row[k]] += (double)val[k] * (double)x[col[k]];
}

void oracle_abs_row_sum(int nnz, int height, const int* row, const int* col, const float* val, const float* x,
                        double* s) {
    for (int i = 0; i < height; ++i) s[i] = 0.0;
    for (int k = 0; k < nnz; ++k) s[row[k]] += fabs((double)val[k] * (double)x[col[k]]);
}

/* dia_spmv (spmv_serial.h:11-34): y += A*x over the aligned length. */
void oracle_dia_spmv(int ndiag, int L, const int* offsets, const float* data, const float* x, int width, int height,
                     float* y) {
    for (int row = 0; row < L && row < height; ++row) {
        float acc = 0.0f;
        for (int j = 0; j < ndiag; ++j) {
            int c = row + offsets[j];
            float xv = (c >= 0 && c < width) ? x[c] : 0.0f;
            acc = acc + xv * data[row + (size_t)j * L];
        }
        y[row] = y[row] + acc;
    }
}

/* bdia_spmv (spmv_serial.h:59-90). */
void oracle_bdia_spmv(int nbands, int L, const int* bptr, const int* offsets, const float* data, const float* x,
                      int width, int height, float* y) {
    for (int row = 0; row < L && row < height; ++row) {
        float acc = 0.0f;
        for (int b = 0; b < nbands; ++b) {
            int off = offsets[b];
            for (int i = bptr[b]; i < bptr[b + 1]; ++i, ++off) {
                int c = row + off;
                float xv = (c >= 0 && c < width) ? x[c] : 0.0f;
                acc = acc + xv * data[row + (size_t)i * L];
            }
        }
        y[row] = y[row] + acc;
    }
}

/* ell_spmv (spmv_serial.h:93-110): slot-major. */
void oracle_ell_spmv(int ell_num, int H, const int* col, const float* data, const float* x, int height, float* y) {
    for (int i = 0; i < ell_num; ++i)
        for (int j = 0; j < height; ++j) y[j] = y[j] + data[(size_t)i * H + j] * x[col[(size_t)i * H + j]];
}

/* sell_spmv (spmv_serial.h:112-135). */
void oracle_sell_spmv(int S, int nslices, const int* slice_ptr, const int* col, const float* data, const float* x,
                      int height, float* y) {
    for (int s = 0; s < nslices; ++s) {
        int start = slice_ptr[s];
        int w = (slice_ptr[s + 1] - start) / S;
        for (int j = 0; j < w; ++j)
            for (int row = s * S; row < (s + 1) * S && row < height; ++row) {
                size_t id = (size_t)start + (row - s * S) + (size_t)j * S;
                y[row] = y[row] + data[id] * x[col[id]];
            }
    }
}

/* csr_spmv (spmv_serial.h:198-217). */
void oracle_csr_spmv(int height, const int* rp, const int* col, const float* data, const float* x, float* y) {
    for (int r = 0; r < height; ++r) {
        float acc = 0.0f;
        for (int k = rp[r]; k < rp[r + 1]; ++k) acc = acc + data[k] * x[col[k]];
        y[r] = y[r] + acc;
    }
}

/* b4csr layout (matrix_storage.h:1872-1895): block i of block row r, value
 * (h, w4*4+w) at data[w4*(N*bh) + h*N + 4*i + w], block column in float4
 * units. */
void oracle_b4csr_spmv(int brows, int bh, int bw, int N, const int* rp, const int* col4, const float* data,
                       const float* x, int width, int height, float* y) {
    int w4n = bw / 4;
    for (int r = 0; r < brows; ++r)
        for (int i = rp[r]; i < rp[r + 1]; ++i)
            for (int h = 0; h < bh; ++h) {
                int row = r * bh + h;
                if (row >= height) continue;
                for (int w4 = 0; w4 < w4n; ++w4)
                    for (int w = 0; w < 4; ++w) {
                        int c = col4[i] * 4 + w4 * 4 + w;
                        if (c >= width) continue;
                        y[row] = y[row] + data[(size_t)w4 * N * bh + (size_t)h * N + (size_t)i * 4 + w] * x[c];
                    }
            }
}

/* b4ell layout (matrix_storage.h:2150-2196): slot j, value (h, w4*4+w) of
 * block row r at data[j*(F*w4n*bh) + h*F + w4*(F*bh) + 4*r + w]. */
void oracle_b4ell_spmv(int brows, int bh, int bw, int slots, int Hc, int F, const int* col4, const float* data,
                       const float* x, int width, int height, float* y) {
    int w4n = bw / 4;
    for (int j = 0; j < slots; ++j)
        for (int r = 0; r < brows; ++r)
            for (int h = 0; h < bh; ++h) {
                int row = r * bh + h;
                if (row >= height) continue;
                for (int w4 = 0; w4 < w4n; ++w4)
                    for (int w = 0; w < 4; ++w) {
                        int c = col4[(size_t)j * Hc + r] * 4 + w4 * 4 + w;
                        if (c >= width) continue;
                        size_t id = (size_t)j * F * w4n * bh + (size_t)h * F + (size_t)w4 * F * bh + (size_t)r * 4 + w;
                        y[row] = y[row] + data[id] * x[c];
                    }
            }
}

/* sbell_spmv (spmv_serial.h:137-182). */
void oracle_sbell_spmv(int S, int nslices, int bh, int bw, const int* slice_ptr, const int* col4,
                       const float* data, const float* x, int width, int height, float* y) {
    int bs = bw * bh, w4n = bw / 4;
    for (int i = 0; i < nslices; ++i) {
        int start = slice_ptr[i];
        int ell = (slice_ptr[i + 1] - start) / S;
        for (int s = 0; s < S; ++s) {
            size_t coloff = (size_t)start + s;
            size_t dataoff = (size_t)start * bs + (size_t)s * 4;
            for (int j = 0; j < ell; ++j) {
                int c = col4[coloff] * 4;
                for (int w4 = 0; w4 < w4n; ++w4) {
                    for (int h = 0; h < bh; ++h) {
                        int row = (i * S + s) * bh + h;
                        if (row < height)
                            for (int w = 0; w < 4; ++w) {
                                if (c + w >= width) break;
                                y[row] = y[row] + data[dataoff + w] * x[c + w];
                            }
                        dataoff += (size_t)S * 4;
                    }
                    c += 4;
                }
                coloff += S;
            }
        }
    }
}

/* native r x c BSR: element-major values data[e*stride + i]. */
void oracle_bsr_spmv(int brows, int bh, int bw, int stride, const int* rp, const int* col, const float* data,
                     const float* x, int width, int height, float* y) {
    for (int r = 0; r < brows; ++r)
        for (int i = rp[r]; i < rp[r + 1]; ++i)
            for (int h = 0; h < bh; ++h) {
                int row = r * bh + h;
                if (row >= height) continue;
                for (int w = 0; w < bw; ++w) {
                    int c = col[i] + w;
                    if (c >= width) continue;
                    y[row] = y[row] + data[(size_t)(h * bw + w) * stride + i] * x[c];
                }
            }
}

/* two_vec_compare (A/src/util.cpp:93-131): Euclidean distance, max
 * |(|a|-|b|)| and max of that over max(|a|,|b|) (sign-blind as in the
 * reference).  out[0..2] = distance, max diff, max ratio. */
void oracle_two_vec_compare(const float* a, const float* b, int n, double* out) {
    double sum = 0.0, maxdiff = 0.0, maxratio = 0.0;
    for (int i = 0; i < n; ++i) {
        double t = (double)(a[i] - b[i]);
        sum += t * t;
        float ta = a[i] < 0 ? -a[i] : a[i];
        float tb = b[i] < 0 ? -b[i] : b[i];
        double diff = (double)ta - (double)tb;
        if (diff < 0) diff = -diff;
        float mx = ta > tb ? ta : tb;
        double ratio = mx > 0 ? diff / mx : 0.0;
        if (diff > maxdiff) maxdiff = diff;
        if (ratio > maxratio) maxratio = ratio;
    }
    out[0] = sqrt(sum);
    out[1] = maxdiff;
    out[2] = maxratio;
}

/* y = A x, CSR rows split over OpenMP threads (dynamic chunks of 1024 rows);
 * each row summed in order in FP32.  Returns the number of threads used. */
int oracle_csr_spmv_omp(int height, const int* rp, const int* col, const float* data, const float* x, float* y) {
    int used = 1;
#pragma omp parallel
    {
#pragma omp single
        used = omp_get_num_threads();
#pragma omp for schedule(dynamic, 1024)
        for (int r = 0; r < height; ++r) {
            float s = 0.0f;
            for (int k = rp[r]; k < rp[r + 1]; ++k) s += data[k] * x[col[k]];
            y[r] = s;
        }
    }
    return used;
}
