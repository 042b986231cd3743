/*
 * clspmv.h -- C ABI of the MI355X-native clSpMV engine (libclspmv.so).
 *
 * This header is the drop-in boundary.  The reference (clSpMV, OpenCL 1.1)
 * exposes C++ templates instantiated for <int,float> plus an OpenCL runtime
 * shim; here every entry point is extern "C" with plain pointers and sizes.
 * Each declaration cites the reference interface it replaces
 * (paths relative to the reference tree, A/ = clSpMV_AMD/).
 *
 * Semantics kept from the reference:
 *   - every SpMV computes y <- A*x + y (kernels read the old y first,
 *     A/kernels/spmv_dia.cl:15, A/src/eval.cpp:40-42);
 *   - struct layouts and field names of A/include/matrix_storage.h:15-365
 *     and A/include/analyze.h:105-144 (int32 indices, float values);
 *   - conversions return the reference's bool as CLSPMV_OK / an error code
 *     and allocate with malloc; the caller releases with clspmv_free_<fmt>.
 * Differences (documented, never silent):
 *   - no exit(1)/assert: every failure is an integer status code;
 *   - OpenCL context/queue/cl_mem become clspmv_ctx_t (device + HIP stream)
 *     and plain device pointers (void*);
 *   - the texture (image2d) kernel variants do not exist on CDNA4; the
 *     use_tex flag is accepted and ignored.
 *
 * Threading: one host thread per context; calls on one context are not
 * thread-safe.  All GPU work of a context is issued on its one stream in
 * order, so Cocktail parts that accumulate into the same y are serialised
 * exactly as on the reference's in-order queue.
 */
#ifndef CLSPMV_H_
#define CLSPMV_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes (replace CHECKERROR/exit(1), A/include/oclcommon.h:10-38,
 * and the bool returns of A/include/matrix_storage.h conversions).     */
#define CLSPMV_OK              0
#define CLSPMV_E_INVALID      -1  /* bad argument / shape                */
#define CLSPMV_E_TOO_LARGE    -2  /* MAX_MEM_OBJ refusal (matrix_storage.h:1863,2141,2350) */
#define CLSPMV_E_PADDING      -3  /* DIA padding >= 2*nnz (matrix_storage.h:972-979) */
#define CLSPMV_E_HIP          -4  /* HIP runtime error                   */
#define CLSPMV_E_NOMEM        -5  /* host or device allocation failed    */
#define CLSPMV_E_IO           -6  /* file cannot be opened / parsed      */
#define CLSPMV_E_SORT         -7  /* sort_coo exceeded MAX_LEVELS (matrix_storage.h:774) */
#define CLSPMV_E_UNSUPPORTED  -8  /* shape/method not provided           */
#define CLSPMV_E_NODEVICE     -9  /* no HIP device visible               */

/* Reference compile-time constants that shape the data layouts
 * (A/include/constant.h:8-53). */
#define CLSPMV_GPU_ALIGNMENT   32
#define CLSPMV_COO_ALIGNMENT   32
#define CLSPMV_WARPSIZE        32
#define CLSPMV_MAX_MEM_OBJ     268435456
#define CLSPMV_MAX_DIA_NUM     128
#define CLSPMV_MAX_BAND_NUM    31
#define CLSPMV_MAX_LEVELS      1000

/* Format ids, numbered like spmv_all's methods (A/src/spmv_all.cpp:196-250). */
#define CLSPMV_FMT_CSR_SCALAR  1
#define CLSPMV_FMT_CSR_VECTOR  2
#define CLSPMV_FMT_BDIA        3
#define CLSPMV_FMT_DIA         4
#define CLSPMV_FMT_ELL         5
#define CLSPMV_FMT_COO         6
#define CLSPMV_FMT_BELL        7
#define CLSPMV_FMT_BCSR        8
#define CLSPMV_FMT_SELL        9
#define CLSPMV_FMT_SBELL      10
#define CLSPMV_FMT_BSR        11  /* native r x c BCSR (e.g. 3x3); no reference counterpart */

/* ------------------------------------------------------------------ */
/* Data model: A/include/matrix_storage.h:15-365 (<int,int,float>).    */
typedef struct { int width; int height; int nnz; } clspmv_matinfo;

typedef struct {                 /* coo_matrix :141-152 */
    clspmv_matinfo matinfo;
    int*   coo_row_id;
    int*   coo_col_id;
    float* coo_data;
} clspmv_coo_matrix;

typedef struct {                 /* csr_matrix :155-166 */
    clspmv_matinfo matinfo;
    int*   csr_row_ptr;
    int*   csr_col_id;
    float* csr_data;
} clspmv_csr_matrix;

typedef struct {                 /* dia_matrix :30-47 */
    clspmv_matinfo matinfo;
    int    dia_num;
    int    dia_length;
    int    dia_length_aligned;
    int*   dia_offsets;
    float* dia_data;
} clspmv_dia_matrix;

typedef struct {                 /* bdia_matrix :72-94 */
    clspmv_matinfo matinfo;
    int    bdia_band_num;
    int    bdia_length;
    int    bdia_length_aligned;
    int*   bdia_offsets;
    int*   bdia_bptr;
    float* bdia_data;
} clspmv_bdia_matrix;

typedef struct {                 /* ell_matrix :96-111 */
    clspmv_matinfo matinfo;
    int    ell_num;
    int    ell_height_aligned;
    int*   ell_col_id;
    float* ell_data;
} clspmv_ell_matrix;

typedef struct {                 /* sell_matrix :122-137 */
    clspmv_matinfo matinfo;
    unsigned int sell_slice_height;
    int    sell_slice_num;
    int*   sell_slice_ptr;
    int*   sell_col_id;
    float* sell_data;
} clspmv_sell_matrix;

typedef struct {                 /* b4csr_matrix :266-287 (GPU BCSR) */
    clspmv_matinfo matinfo;
    unsigned int b4csr_bwidth;
    unsigned int b4csr_bheight;
    int    b4csr_row_num;
    int    b4csr_aligned_size;
    int    b4csr_block_num;
    int*   b4csr_row_ptr;
    int*   b4csr_col_id;
    float* b4csr_data;
} clspmv_b4csr_matrix;

typedef struct {                 /* b4ell_matrix :311-332 (GPU BELL) */
    clspmv_matinfo matinfo;
    unsigned int b4ell_bwidth;
    unsigned int b4ell_bheight;
    int    b4ell_row_num;
    int    b4ell_block_num;
    int    b4ell_height_aligned;
    int    b4ell_float4_aligned;
    int*   b4ell_col_id;
    float* b4ell_data;
} clspmv_b4ell_matrix;

typedef struct {                 /* sbell_matrix :344-365 */
    clspmv_matinfo matinfo;
    unsigned int sbell_bwidth;
    unsigned int sbell_bheight;
    unsigned int sbell_slice_height;
    int    sbell_slice_num;
    int    sbell_row_num;
    int*   sbell_slice_ptr;
    int*   sbell_col_id;
    float* sbell_data;
} clspmv_sbell_matrix;

/* Native r x c block CSR (no reference counterpart; the reference only has
 * {1,2,4,8}x{4,8} shapes, A/src/analyze.cpp:545-546).  Blocks of one block
 * row are contiguous; values are stored element-major:
 *   bsr_data[e * bsr_block_stride + i],  e = h*bwidth + w,  i = block id,
 * bsr_block_stride = align(block_num, 64), so lanes that read consecutive
 * blocks issue one coalesced stream per element slot.                  */
typedef struct {
    clspmv_matinfo matinfo;
    unsigned int bsr_bwidth;
    unsigned int bsr_bheight;
    int    bsr_row_num;          /* ceil(height / bheight) */
    int    bsr_block_num;
    int    bsr_block_stride;
    int*   bsr_row_ptr;          /* size bsr_row_num + 1 */
    int*   bsr_col_id;           /* first column of each block (element units) */
    float* bsr_data;             /* size bsr_block_stride * bwidth * bheight */
} clspmv_bsr_matrix;

/* ------------------------------------------------------------------ */
/* Host data model entry points.                                        */

/* init_xxx / free_xxx (matrix_storage.h:370-652): zero the struct or free
 * its arrays (pointers are reset to NULL). */
void clspmv_init_coo(clspmv_coo_matrix* m);
void clspmv_free_coo(clspmv_coo_matrix* m);
void clspmv_free_csr(clspmv_csr_matrix* m);
void clspmv_free_dia(clspmv_dia_matrix* m);
void clspmv_free_bdia(clspmv_bdia_matrix* m);
void clspmv_free_ell(clspmv_ell_matrix* m);
void clspmv_free_sell(clspmv_sell_matrix* m);
void clspmv_free_b4csr(clspmv_b4csr_matrix* m);
void clspmv_free_b4ell(clspmv_b4ell_matrix* m);
void clspmv_free_sbell(clspmv_sbell_matrix* m);
void clspmv_free_bsr(clspmv_bsr_matrix* m);

/* Allocate a COO of the given shape (arrays malloc'd, uninitialised). */
int clspmv_alloc_coo(clspmv_coo_matrix* m, int width, int height, int nnz);

/* if_sorted_coo (matrix_storage.h:731-741): 1 if row-major sorted. */
int clspmv_if_sorted_coo(const clspmv_coo_matrix* m);
/* sort_coo (matrix_storage.h:744-815): in-place, same element order as the
 * reference for equal keys; CLSPMV_E_SORT past MAX_LEVELS. */
int clspmv_sort_coo(clspmv_coo_matrix* m);
/* aligned_length / findPaddedSize (matrix_storage.h:817, util.cpp:17). */
int clspmv_aligned_length(int length, int alignment);

/* Conversions.  Each sorts an unsorted source in place first, as the
 * reference does.                                                      */
int clspmv_coo2csr(clspmv_coo_matrix* src, clspmv_csr_matrix* dst);                          /* :1268 */
int clspmv_coo2dia(clspmv_coo_matrix* src, clspmv_dia_matrix* dst, int alignment);          /* :936  */
int clspmv_coo2bdia(clspmv_coo_matrix* src, clspmv_bdia_matrix* dst, int alignment);        /* :1184 */
int clspmv_coo2ell(clspmv_coo_matrix* src, clspmv_ell_matrix* dst, int alignment, int ellnum); /* :1340 */
int clspmv_coo2sell(clspmv_coo_matrix* src, clspmv_sell_matrix* dst, unsigned int slice_height); /* :1397 */
int clspmv_coo2b4csr(clspmv_coo_matrix* src, clspmv_b4csr_matrix* dst,
                     unsigned int bwidth, unsigned int bheight, int alignment);             /* :1740 */
int clspmv_coo2b4ell(clspmv_coo_matrix* src, clspmv_b4ell_matrix* dst,
                     unsigned int bwidth, unsigned int bheight, int alignment, int ellnum); /* :2009 */
int clspmv_coo2sbell(clspmv_coo_matrix* src, clspmv_sbell_matrix* dst,
                     unsigned int bwidth, unsigned int bheight, unsigned int slice_height); /* :2202 */
/* Native r x c BCSR (build extension, any 1<=r,c<=8). */
int clspmv_coo2bsr(clspmv_coo_matrix* src, clspmv_bsr_matrix* dst,
                   unsigned int bwidth, unsigned int bheight);
/* pad_csr (A/src/util.cpp:31-72). */
int clspmv_pad_csr(const clspmv_csr_matrix* src, clspmv_csr_matrix* dst, int alignment);

/* ReadMMF (A/src/fileio.cpp:10-109), same quirks: the 4th header token
 * "real" selects real values, anything else reads 1.0 without consuming a
 * value token; "symmetric" mirrors off-diagonal entries; duplicates kept. */
int clspmv_read_mmf(const char* filename, clspmv_coo_matrix* m);
/* Matrix Market writer (real general, 1-based, %.9g values). */
int clspmv_write_mmf(const char* filename, const clspmv_coo_matrix* m);
/* write_coo (matrix_storage.h:656-671): width,height,nnz,rows,cols,data. */
int clspmv_write_coo_bin(const char* filename, const clspmv_coo_matrix* m);
/* Working replacement of the broken read_coo (matrix_storage.h:673-696). */
int clspmv_read_coo_bin(const char* filename, clspmv_coo_matrix* m);

/* ------------------------------------------------------------------ */
/* Seeded synthetic generators (SplitMix64-seeded xoshiro256**; no
 * duplicate (row, col); output row-major sorted).                      */
int clspmv_gen_laplacian2d(clspmv_coo_matrix* m, int nx, int ny);
int clspmv_gen_stencil27(clspmv_coo_matrix* m, int n, uint64_t seed);
int clspmv_gen_uniform_rows(clspmv_coo_matrix* m, int rows, int per_row, uint64_t seed);
int clspmv_gen_zipf_rows(clspmv_coo_matrix* m, int rows, double exponent, double mean,
                         int max_len, uint64_t seed);
int clspmv_gen_block_rows(clspmv_coo_matrix* m, int block_rows, int bdim, int blocks_per_row,
                          int band, uint64_t seed);
int clspmv_gen_cocktail_composite(clspmv_coo_matrix* m, int rows, uint64_t seed);
/* x_i ~ U[-1,1] from seed. */
int clspmv_gen_vector(float* x, int n, uint64_t seed);

/* ------------------------------------------------------------------ */
/* HIP runtime: replaces initialization()/freeObjects()
 * (A/src/oclcommon.cpp:259-422, A/include/oclcommon.h:45-46).          */
typedef struct clspmv_ctx_s* clspmv_ctx_t;

int clspmv_init(int device, clspmv_ctx_t* ctx);
int clspmv_finalize(clspmv_ctx_t ctx);
/* Use an external hipStream_t (NULL = the context's own stream). */
int clspmv_set_stream(clspmv_ctx_t ctx, void* stream);
int clspmv_synchronize(clspmv_ctx_t ctx);
const char* clspmv_last_error(void);

/* Device buffers (ALLOCATE_GPU_READ, oclcommon.h:20-38). */
int clspmv_device_alloc(clspmv_ctx_t ctx, size_t bytes, void** dptr);
int clspmv_device_free(clspmv_ctx_t ctx, void* dptr);
int clspmv_memcpy_h2d(clspmv_ctx_t ctx, void* dst, const void* src, size_t bytes);
int clspmv_memcpy_d2h(clspmv_ctx_t ctx, void* dst, const void* src, size_t bytes);
int clspmv_memcpy_d2d(clspmv_ctx_t ctx, void* dst, const void* src, size_t bytes);

/* HIP-event timing on the context stream (replaces gettimeofday around
 * clFinish, A/src/util.cpp:10-15). */
int clspmv_event_elapsed_begin(clspmv_ctx_t ctx);
int clspmv_event_elapsed_end(clspmv_ctx_t ctx, float* ms);

/* ------------------------------------------------------------------ */
/* Single-format device parts: the device half of spmv_<fmt>_ocl
 * (e.g. A/src/spmv_ell.cpp:17-300).  A part owns the uploaded format
 * arrays and the chosen kernel variant; clspmv_part_spmv enqueues
 * ntimes * (y += A*x) on the context stream.  d_x must hold
 * matinfo.width floats, d_y matinfo.height floats (device pointers,
 * 16-byte aligned).                                                    */
typedef struct clspmv_part_s* clspmv_part_t;

typedef struct {
    int    format;         /* CLSPMV_FMT_* */
    int    method;         /* kernel variant actually used */
    int    rows, cols;
    long long nnz;         /* true nonzeros (FLOPs = 2*nnz) */
    long long stored_bytes;/* bytes of every uploaded format array */
    long long alg_bytes;   /* stored_bytes + 4*cols + 8*rows (x read, y read+write) */
    int    launches;       /* kernels per SpMV */
    char   kernel[64];     /* name of the dominant kernel */
    /* Mixed-precision parts: the largest relative error of a stored value
     * against the FP32 value it encodes, max |dec(enc(a)) - a| / |a| over
     * every nonzero a (a zero stored for a nonzero a counts 1).  It bounds
     * the componentwise error of y = A x from the encoding, for any x:
     * |sum_j (dec(a_ij) - a_ij) x_j| <= codec_bound * sum_j |a_ij x_j|.
     * 0 for FP32 parts. */
    double codec_bound;
} clspmv_part_info;

int clspmv_part_csr_scalar(clspmv_ctx_t ctx, const clspmv_csr_matrix* m, int method, clspmv_part_t* p);
int clspmv_part_csr_vector(clspmv_ctx_t ctx, const clspmv_csr_matrix* m, int method, clspmv_part_t* p);
int clspmv_part_dia(clspmv_ctx_t ctx, const clspmv_dia_matrix* m, int method, clspmv_part_t* p);
int clspmv_part_bdia(clspmv_ctx_t ctx, const clspmv_bdia_matrix* m, int method, clspmv_part_t* p);
int clspmv_part_ell(clspmv_ctx_t ctx, const clspmv_ell_matrix* m, int method, clspmv_part_t* p);
int clspmv_part_sell(clspmv_ctx_t ctx, const clspmv_sell_matrix* m, int method, clspmv_part_t* p);
int clspmv_part_coo(clspmv_ctx_t ctx, const clspmv_coo_matrix* m, int method, clspmv_part_t* p);
int clspmv_part_bell(clspmv_ctx_t ctx, const clspmv_b4ell_matrix* m, int method, clspmv_part_t* p);
int clspmv_part_sbell(clspmv_ctx_t ctx, const clspmv_sbell_matrix* m, int method, clspmv_part_t* p);
int clspmv_part_bcsr(clspmv_ctx_t ctx, const clspmv_b4csr_matrix* m, int method, clspmv_part_t* p);
int clspmv_part_bsr(clspmv_ctx_t ctx, const clspmv_bsr_matrix* m, int method, clspmv_part_t* p);

/* Kernel variants (`method`) per part:
 *   ELL      0 one row per lane, 1 four rows per lane, 2 column-sliced,
 *            4 LDS x-tiled
 *   SELL     0 one row per lane, 2 column-sliced, 4 LDS x-tiled
 *   CSR vec  0 L lanes per row, 1 adaptive row blocks, 2 column-sliced,
 *            4 LDS x-tiled
 *   COO      0 segmented reduction, 2 column-sliced, 4 LDS x-tiled
 *   DIA      0 one row per lane, 1 four rows per lane (reference methods
 *            2 / 3 map to 0 / 1), 4 two to eight lanes per row splitting
 *            the diagonals -- all of a row's loads in one round, for small
 *            matrices (one row per lane at <= 8 diagonals)
 *   BDIA     0 direct x reads, 1 LDS-staged x band
 *   BSR      0 sliced (64 block rows per slice; for r x r blocks with
 *            2 <= r <= 4, block columns stored as int16 steps within a
 *            block row when every step fits, kernel name suffix "_d16";
 *            with a wide band, x windows of 1024 block rows staged in
 *            LDS, suffix "_xw"), 1 (or L = 2..16) L lanes per block row
 * "Column-sliced" cuts the columns into slices whose part of x (<= 2 MiB by
 * default) stays in one XCD's L2; one launch per slice adds into y.
 * "LDS x-tiled" gives each 1024-thread workgroup a block of <= 32768 rows
 * whose partial sums stay in LDS and streams x through LDS in 8064-column
 * slices, a block's slices split into up to 8 column groups (for wide x with
 * scattered columns; needs width >= 4).  The
 * reference's method ids (e.g. texture variants) map onto these. */
/* Host arrays of the LDS x-tiled layout a CSR part builds (method 4), as
 * they are uploaded, before the FP32 interleave: no counterpart in the
 * reference; for tests and for callers that inspect the layout (no device
 * work; the options xtile_blocks and xtile_max_groups apply).  Row block b
 * holds rows blk_row[b] .. blk_row[b+1]-1; workgroup w = b * groups + g
 * (column group g) holds tiles blk_tile[w] .. blk_tile[w+1]-1; tile t holds
 * chunks tile_chunk[t] .. tile_chunk[t+1]-1 of 128 entries and has the slice
 * word tile_slice[t] = (column slice << 2) | 2 if one of its chunks has the
 * carry bit | 1 if it continues the previous tile's slice.  Entry j of the
 * chunks: idx[j] = (row in block << 16) | column in the 8064-column slice
 * (bits 0-12), val[j] its value; padding entries have row 32768 and value
 * 0.  Bits 13-15 of a chunk's first two entries hold its flags: entry 0 bit
 * 15 carry (the chunk's first row continues the previous chunk of the
 * tile), bit 14 fast (no row crosses a lane boundary: entries 2l-1 and 2l
 * hold different rows or padding), bit 13 a row continues from lane 31 into
 * 32; entry 1 bits 13-14 the in-row scan steps - 1, bit 15 a row continues
 * from lane 15 into 16 or from 47 into 48 (lane l holds entries 2l, 2l+1).
 * Free with clspmv_free_xtile_layout. */
typedef struct {
    int nblocks, groups, ntiles, nchunks;
    int* blk_row;          /* nblocks + 1 */
    int* blk_tile;         /* nblocks * groups + 1 */
    int* tile_slice;       /* ntiles */
    int* tile_chunk;       /* ntiles + 1 */
    unsigned* idx;         /* nchunks * 128 */
    float* val;            /* nchunks * 128 */
} clspmv_xtile_layout;
int clspmv_xtile_layout_csr(const clspmv_csr_matrix* m, clspmv_xtile_layout* out);
void clspmv_free_xtile_layout(clspmv_xtile_layout* l);

/* Mixed-precision parts (no counterpart in the reference; for callers that
 * trade value precision for bytes per nonzero): values stored as
 *   CLSPMV_VAL_BF16X2    hi + lo, two BF16 halves (4 B, ~16 significant bits)
 *   CLSPMV_VAL_FP16_ROW  FP16 times one power-of-two FP32 scale per row
 *                        (block row) (2 B)
 *   CLSPMV_VAL_FP8_SLICE FP8 E4M3 (OCP) times one FP32 scale per slice: 256
 *                        rows of a diagonal, 64 block rows of a block
 *                        element, or one 128-entry chunk of the x-tiled
 *                        layout (1 B)
 *   CLSPMV_VAL_I16_ROW   signed 16-bit fixed point times one FP32 scale per
 *                        row (block row), the row's largest magnitude over
 *                        32767: every value within half a step of its FP32
 *                        value (2 B; finite values only, else
 *                        CLSPMV_E_INVALID)
 * accumulated in FP32.  DIA: dia_length_aligned must be a multiple of 256
 * (clspmv_coo2dia with alignment 256).  BSR: width >= block width.  BSR r x r (2 <= r <= 4), sliced
 * layout; delta_cols != 0 stores block columns as int16 steps within a block
 * row when every step fits (else int32; info.method tells: 16 + codec, + 8
 * with int16 columns).  CSR: the LDS x-tiled layout (whose entries pack row
 * and column in 4 bytes) with encoded values; info.method = 16 + codec. */
#define CLSPMV_VAL_BF16X2     1
#define CLSPMV_VAL_FP16_ROW   2
#define CLSPMV_VAL_FP8_SLICE  3
#define CLSPMV_VAL_I16_ROW    4
int clspmv_part_dia_mixed(clspmv_ctx_t ctx, const clspmv_dia_matrix* m, int codec, clspmv_part_t* p);
int clspmv_part_bsr_mixed(clspmv_ctx_t ctx, const clspmv_bsr_matrix* m, int codec, int delta_cols, clspmv_part_t* p);
int clspmv_part_csr_mixed(clspmv_ctx_t ctx, const clspmv_csr_matrix* m, int codec, clspmv_part_t* p);
int clspmv_part_spmv(clspmv_part_t p, const float* d_x, float* d_y, int ntimes);
/* Process-wide tuning knobs (no reference counterpart); each affects parts
 * created afterwards:
 *   "col_slice_bytes"   x bytes per column slice of the column-sliced
 *                       layouts (default 2 MiB);
 *   "bsr_lds_window"    1 (default): sliced BSR parts (FP32 and mixed) stage
 *                       each 1024-block-row workgroup's x window in LDS
 *                       when every window fits (<= 40960 columns), their
 *                       mean exceeds 2048 bw columns and there are >= 128
 *                       windows; 2: whenever every window fits; 0: never;
 *   "xtile_blocks"      target workgroups of the x-tiled layout (default
 *                       256, one per CU);
 *   "xtile_max_groups"  column groups per row block of the x-tiled layout
 *                       at most, 1..8 (default 8; with > 1 group a
 *                       partial-sum pass adds the groups' sums to y);
 *   "xtile_scan_prio"   -1 (default): in the x-tiled kernel a wave on a
 *                       scan chunk (one where a row crosses a lane) raises
 *                       its issue priority for that chunk when at least half
 *                       the layout's chunks are fast ones; 0: never; 1:
 *                       always (the results are the same bits either way);
 *   "cocktail_merge_flat"  1 (default): clspmv_init_mat_kernels runs the
 *                       flat parts (ELL, SELL, CSR, COO) of a partition as
 *                       one merged row-list part; 0: one part each;
 *   "cocktail_split_local" 1 (default): the merged flat rows are cut into
 *                       row runs by class (scattered, CSR row blocks, SELL,
 *                       DIA, sliced BSR; classified on blocks of 1024-32768
 *                       rows, boundaries placed to 128 rows), one part per
 *                       run; 0: one part;
 *   "cocktail_value_codec" 0 (default) or a CLSPMV_VAL_* codec for the row-run
 *                       parts that have a mixed-precision variant (x-tiled,
 *                       sliced BSR); the others stay FP32;
 *   "cocktail_codec_tol" error guard of that codec, in units of 1e-9
 *                       (default 100000 = 1e-4; 0 = no guard): a run part
 *                       whose codec_bound (clspmv_part_info) exceeds it is
 *                       built in FP32 instead;
 *   "cocktail_graph"    0 (default): do_spmv and power_iterate launch each
 *                       part directly; 1: they replay captured hipGraphs;
 *   "cocktail_streams"  1 (default): do_spmv runs the row runs of a plan
 *                       side by side on up to four streams; 0: one stream;
 *   "cocktail_small_plan" entries (default 4194304) below which merged flat
 *                       rows that fall into several runs run as one part of
 *                       adaptive row blocks; 0: never. */
int clspmv_set_option(const char* name, long long value);
int clspmv_part_get_info(clspmv_part_t p, clspmv_part_info* info);
/* LDS x-tiled parts (no reference counterpart): tiles per SpMV (one x
 * slice of one row block's column group each) and bytes of x staged into
 * LDS per SpMV (one slice per tile that starts a slice); 0 for other parts. */
int clspmv_part_xtile_stats(clspmv_part_t p, long long* tiles, long long* staged_x_bytes);
int clspmv_part_free(clspmv_part_t p);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc): y += A x for a COO matrix held in the
 * caller's device arrays -- sorted by row, indices in range, row/col/val
 * 16-byte aligned (as clspmv_part_coo requires and checks at upload) --
 * with the kernels of clspmv_part_coo method 0.  It allocates nothing: the
 * chunk carries live in a static device buffer of the library (7.5 MiB per
 * device, held until the process ends), so nnz <= 167,772,160 (else
 * CLSPMV_E_TOO_LARGE).  Every call on one device shares that buffer: calls
 * that overlap on one device (other streams, other threads) race.
 * Asynchronous on the context stream. */
int clspmv_coo_spmv_noalloc(clspmv_ctx_t ctx, int rows, int nnz, const int* d_row, const int* d_col,
                            const float* d_val, const float* d_x, float* d_y);

/* ------------------------------------------------------------------ */
/* Bandwidth / launch-overhead probes: bandwidth_test and bench_overhead
 * (A/src/mem_bandwidth.cpp:14-579, A/src/bench_overhead.cpp:14-184). */
/* Copy kernel of `bytes` (a multiple of 16) run ntimes: read + write GB/s.
 * The copy is then checked word by word; CLSPMV_E_HIP if dst != src. */
int clspmv_probe_copy(clspmv_ctx_t ctx, const void* d_src, void* d_dst, size_t bytes, int ntimes, float* gbps);
int clspmv_probe_launch(clspmv_ctx_t ctx, int ntimes, float* usec_per_launch);
/* Empty kernel of `groups` x `threads` launched ntimes back to back: the
 * per-launch time of the overhead table (A/src/bench_overhead.cpp:117-165). */
int clspmv_probe_launch_grid(clspmv_ctx_t ctx, int groups, int threads, int ntimes, float* usec_per_launch);
/* A captured hipGraph of `nodes` empty kernels launched ntimes back to
 * back: microseconds per graph launch (the graph column of the overhead
 * table; the Cocktail driver replays its passes from such graphs). */
int clspmv_probe_graph_launch(clspmv_ctx_t ctx, int nodes, int ntimes, float* usec_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* CLSPMV_H_ */
