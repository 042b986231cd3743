// Device-side views of the uploaded formats and the host launch wrappers of
// each kernel family.  All launches are asynchronous on the given stream and
// return the hipError_t of the launch as int.
#pragma once
#include <hip/hip_runtime.h>

namespace clspmv {

struct DevEll {
    const int* col;
    const float* val;
    int H, ell_num, rows;
};
struct DevSell {
    const int* slice_ptr;
    const int* col;
    const float* val;
    int S, rows;
};
struct DevCsr {
    const int* row_ptr;
    const int* col;
    const float* val;
    int rows;
};
// Row blocks of the adaptive CSR kernel: {r0, r1 or -(slot+1), n0, n1}.
struct DevCsrAdaptive {
    DevCsr csr;
    const int4* blocks;
    int nblocks;
    double* carry;            // one partial per long-row chunk
    const int* long_rows;     // row id of every long row
    const int* long_first;    // first carry slot of every long row (size n_long + 1)
    int n_long;
};
// Column-sliced CSR: the columns are cut into slices whose part of x fits in
// one XCD's L2; slice s is processed by launch s (launches are ordered on the
// stream and each adds into y).  Entries are stored slice-major; per slice a
// row's entry count is one byte (rows with more than 64 entries in a slice
// get VECTOR/LONG blocks whose descriptor carries the range).
constexpr int kMaxColSlices = 64;
struct DevCsrSliced {
    const int* col;
    const float* val;
    const unsigned char* counts;  // nslices * rows
    const int4* blocks;           // all slices' blocks, slice-major
    int slice_block_begin[kMaxColSlices + 1];  // host-side block ranges per slice
    int nslices, rows;
    double* carry;
    const int* long_rows;
    const int* long_first;
    int n_long;
};
struct DevCoo {
    const int* row;
    const int* col;
    const float* val;
    int nnz_padded;           // multiple of the chunk size
    int chunks;
    int rows;
    int* carry_row;           // 2 per chunk
    double* carry_val;        // 2 per chunk
};
struct DevDia {
    const int* offsets;
    const float* data;
    int L, ndiag, rows, cols;
};
struct DevBdia {
    const int* bptr;
    const int* offsets;
    const float* data;
    int L, nbands, rows, cols;
};
struct DevBell {  // b4ell
    const int* col;
    const float* data;
    int Hc, F4, slots, brows, bh, w4n, rows, cols;
};
struct DevSbell {
    const int* slice_ptr;
    const int* col;
    const float* data;
    int S, brows, bh, w4n, rows, cols;
};
struct DevBcsr {  // b4csr
    const int* row_ptr;
    const int* col;
    const float* data;
    int N4, brows, bh, w4n, rows, cols, lanes;
};
struct DevBsr {
    const int* row_ptr;   // lanes > 0: block-CSR row pointer; lanes == 0: slice pointer
    const int* col;
    const float* data;
    int stride, brows, bh, bw, rows, cols, lanes;
};
constexpr int kBsrSlice = 64;  // block rows per slice of the sliced BSR layout

int launch_ell(const DevEll& m, int method, const float* x, float* y, hipStream_t st);
int launch_sell(const DevSell& m, const float* x, float* y, hipStream_t st);
int launch_csr_scalar(const DevCsr& m, const float* x, float* y, hipStream_t st);
int launch_csr_vector(const DevCsr& m, int lanes, const float* x, float* y, hipStream_t st);
int launch_csr_adaptive(const DevCsrAdaptive& m, const float* x, float* y, hipStream_t st);
int launch_csr_sliced(const DevCsrSliced& m, const float* x, float* y, hipStream_t st);
int launch_coo(const DevCoo& m, const float* x, float* y, hipStream_t st);
int launch_dia(const DevDia& m, int method, const float* x, float* y, hipStream_t st);
int launch_bdia(const DevBdia& m, int method, const float* x, float* y, hipStream_t st);
int launch_bell(const DevBell& m, const float* x, float* y, hipStream_t st);
int launch_sbell(const DevSbell& m, const float* x, float* y, hipStream_t st);
int launch_bcsr(const DevBcsr& m, const float* x, float* y, hipStream_t st);
int launch_bsr(const DevBsr& m, const float* x, float* y, hipStream_t st);
int launch_copy_probe(const float4* src, float4* dst, size_t n4, hipStream_t st);
int launch_empty(hipStream_t st);

// Constants of the adaptive CSR and COO kernels, shared with the host
// planners that build their work lists.
constexpr int kCsrStreamPerLane = 8;                    // nnz per lane in a stream block
constexpr int kCsrStreamMaxNnz = 64 * kCsrStreamPerLane;  // 512
constexpr int kCsrStreamMaxRowLen = 64;                 // longer rows get their own wave
constexpr int kCsrLongChunk = 8192;                     // nnz per wave for very long rows
constexpr int kCooPerLane = 8;
constexpr int kCooChunk = 64 * kCooPerLane;             // 512 nnz per wave

}  // namespace clspmv
