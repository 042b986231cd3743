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
    int nnz;                  // entries the arrays hold (nnz_padded, or fewer: the last chunk is padded on the fly)
    int chunks;
    int rows;
    int* carry_row;           // 2 per chunk
    double* carry_val;        // 2 per chunk
};
struct DevDia {
    const int* offsets;
    const float* data;
    int L, ndiag, rows, cols;
    int offmin, offmax;  // smallest and largest diagonal offset
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
    const int2* win;      // sliced layout: per-workgroup x windows (LDS-windowed kernel), or null
};
constexpr int kBsrSlice = 64;  // block rows per slice of the sliced BSR layout
// LDS-windowed sliced BSR: workgroups of kBsrWinRows block rows (one per
// thread) stage x[w.x .. w.x + w.y) -- every column their blocks touch --
// in LDS once, w.y <= kBsrWinFloats (160 KiB).
constexpr int kBsrWinRows = 1024;
constexpr int kBsrWinFloats = 40960;

// Mixed-precision parts (k_mixed.hip, k_xtile.hip): value codecs, as CLSPMV_VAL_* in the ABI.
constexpr int kCodecBf16x2 = 1;
constexpr int kCodecFp16Row = 2;
constexpr int kCodecFp8Slice = 3;
constexpr int kCodecI16Row = 4;
// Codecs whose values are divided by one scale per row (per block row for
// BSR); the kernels multiply the row sums by it.
constexpr bool codec_row_scaled(int c) { return c == kCodecFp16Row || c == kCodecI16Row; }

// LDS x-tiled layout (k_xtile.hip): row blocks of <= kXtRows rows, column
// slices of kXtCols columns, entries packed (row_in_block << 16 |
// col_in_slice) in chunks of kXtChunk; padding entries carry row kXtPadRow.
// The column halves of a chunk's first two entries carry the chunk's flags:
// entry 0 kXtCarryBit when its first row continues the previous chunk of
// the same tile, kXtFastBit when no row crosses one of its lane boundaries,
// and the scan plan of the other chunks -- entry 1 bits 13-14: in-row steps
// - 1, entry 1 kXtBc15Bit / entry 0 kXtBc31Bit: a row continues from lane 15
// into 16 or 47 into 48 / from lane 31 into 32 (k_xtile.hip).
constexpr int kXtThreads = 1024;
constexpr int kXtWaves = kXtThreads / 64;
// LDS: y partial sums of kXtRows rows (128 KiB) + 64 dummy slots (256 B) +
// one x slice of kXtCols columns (31.5 KiB) + 2 x 16 carry slots (256 B) =
// 160 KiB, all of it.
constexpr int kXtRows = 32768;
constexpr int kXtCols = 8064;
constexpr int kXtX4 = kXtCols / 4;                                  // float4 per slice
constexpr int kXtXLoads = (kXtX4 + kXtThreads - 1) / kXtThreads;    // float4 per thread
constexpr int kXtChunk = 128;
constexpr int kXtRounds = 2;                       // chunk rounds per tile held in registers
constexpr int kXtTileChunks = kXtRounds * kXtWaves;  // a slice with more chunks is split
constexpr int kXtDepth = 3;                        // tiles prefetched ahead
// Padding entries' row: the first of 64 per-lane dummy slots that follow a
// block's kXtRows sums in LDS (written, never read).
constexpr unsigned kXtPadRow = kXtRows;
constexpr unsigned kXtCarryBit = 0x8000u;  // entry 0
constexpr unsigned kXtFastBit = 0x4000u;   // entry 0
constexpr unsigned kXtBc31Bit = 0x2000u;   // entry 0
constexpr unsigned kXtBc15Bit = 0x8000u;   // entry 1
constexpr int kXtStepsShift = 13;          // entry 1, 2 bits
constexpr unsigned kXtColMask = 0x1FFFu;
static_assert(kXtCols <= (int)kXtColMask + 1, "packed entry fields");
constexpr int kXtMaxGroups = 8;    // column groups per row block at most
struct DevXtile {
    const int* blk_row;     // nblocks + 1 row boundaries
    const int* blk_tile;    // nblocks + 1 tile ranges
    const int* tile_slice;  // (column slice << 2) | 2 if a chunk of the tile has the carry bit
                            //   | 1 if the tile continues the previous tile's slice
    const int* tile_chunk;  // ntiles + 1 chunk ranges, at most kXtTileChunks per tile
    const unsigned* idx;    // per chunk and lane: FP32, BF16 hi+lo {entry, entry, value, value};
                            //   FP16, int16 {entry, entry, two 2-byte values}; FP8 {entry, entry}
    const void* val;        // FP8 values (the other codecs: idx)
    const float* vscale;    // FP16: rows row scales; FP8: nchunks chunk scales
    float* partial;         // groups * rows partial sums when groups > 1
    int nblocks;            // workgroups = row blocks * groups
    int groups;             // column groups (1..8); workgroup b works on group b % groups
    int ntiles, nchunks, cols, rows;
    int codec;              // 0 (FP32) or a kCodec* value
    int scan_prio;          // 1: waves on scan chunks raise their issue priority (k_xtile.hip: xt_chunk)
};

struct DevDiaMixed {
    const int* offsets;
    const void* data;     // ndiag * L encoded values
    const float* scale;   // FP16: L row scales; FP8: ndiag * nsl scales of 256-row slices
    int L, ndiag, nsl, rows, cols, codec;
    int offmin, offmax;   // smallest and largest diagonal offset
};
struct DevBsrMixed {
    const int* sp;        // slice pointer of the sliced layout
    const void* col;      // int32 element columns, or int16 deltas
    const int* base;      // per block row: column the first delta is added to
    const void* data;     // bh*bw * stride encoded values, per 64-slot step the bh*bw elements' 64 values in turn
    const float* scale;   // FP16: one per block row; FP8: bh*bw * nsl
    int stride, brows, nsl, bh, bw, rows, cols, codec, delta;
    const int2* win;      // per-workgroup x windows (LDS-windowed kernel), or null
};

// Batched multi-right-hand-side SpMV (k_spmm.hip).
constexpr int kSpmmStageNnz = 512;    // entries of a STREAM item (staged in LDS)
constexpr int kSpmmLongChunk = 1024;  // entries of a LONG item
struct SpmmMat {
    const int* row_ptr;  // CSR: rows + 1; BSR: block rows + 1
    const int* col;      // CSR: column; BSR: first column of each block
    const float* val;    // CSR: value; BSR: bdim x bdim per block, row-major
    const float* x;      // cols x k, row-major
    float* y;            // rows x k, row-major
    int rows, k;         // k in {1, 4, 16}
    int bdim;            // 0: CSR; 2..4: BSR with bdim x bdim blocks (k = 4 or 16)
    int cols;
};
struct DevSpmm {
    const SpmmMat* mats;
    const int4* items;   // STREAM {mat, r0, r1, 0} / LONG {mat, row, -(slot+1), chunk} /
                         // BLOCK {mat, block row 0, block row 1, 0} of a BSR matrix
    const int4* longs;   // {mat, row, first carry slot, chunks}
    float* carry;        // 16 floats per LONG item
    int n_items, n_long, grid;
};

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
int launch_xtile(const DevXtile& m, const float* x, float* y, hipStream_t st);
int launch_spmm(const DevSpmm& b, hipStream_t st);
int launch_dia_mixed(const DevDiaMixed& m, const float* x, float* y, hipStream_t st);
int launch_bsr_mixed(const DevBsrMixed& m, const float* x, float* y, hipStream_t st);
int launch_copy_probe(const float4* src, float4* dst, size_t n4, hipStream_t st);
int launch_compare_probe(const void* a, const void* b, size_t n4, unsigned long long* count, hipStream_t st);
int launch_empty(hipStream_t st, int groups = 1, int threads = 64);

// Constants of the adaptive CSR and COO kernels, shared with the host
// planners that build their work lists.
constexpr int kCsrStreamPerLane = 8;                    // nnz per lane in a stream block
constexpr int kCsrStreamMaxNnz = 64 * kCsrStreamPerLane;  // 512
constexpr int kCsrStreamMaxRowLen = 64;                 // longer rows get their own wave
// nnz per wave for long rows: a wave takes 256 entries per step, so a chunk
// is 4 dependent gather steps; 8192-entry chunks made one wave's latency
// chain (~40 us) the kernel's tail on Zipf rows of a small matrix
constexpr int kCsrLongChunk = 1024;
constexpr int kCooPerLane = 8;
constexpr int kCooChunk = 64 * kCooPerLane;             // 512 nnz per wave

}  // namespace clspmv
