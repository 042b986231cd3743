// HIP runtime behind the C ABI: contexts, device buffers, per-format device
// parts and the probes.  Replaces the reference's OpenCL shim
// (A/src/oclcommon.cpp:259-466) and the device halves of its single-format
// drivers (e.g. A/src/spmv_ell.cpp:17-300).
//
// Every part validates its index arrays on the host before upload, so a
// kernel can never read outside x or the format arrays, and every y write is
// guarded by the matrix height.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "clspmv.h"
#include "host/codec.hpp"
#include "host/internal.hpp"
#include "kernels/launch.hpp"
#include "runtime.hpp"

using namespace clspmv;

namespace clspmv {

int hip_status(hipError_t e, const char* what) {
    if (e == hipSuccess) return CLSPMV_OK;
    set_error("%s: %s", what, hipGetErrorString(e));
    return CLSPMV_E_HIP;
}

int DevBuf::upload(clspmv_ctx_t ctx, const void* host, size_t bytes) {
    size = bytes;
    hipError_t e = hipMalloc(&ptr, bytes ? bytes : 4);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_status(e, "hipMalloc");
    }
    if (bytes && host) {
        e = hipMemcpyAsync(ptr, host, bytes, hipMemcpyHostToDevice, ctx->stream());
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream());
        if (e != hipSuccess) return hip_status(e, "hipMemcpy H2D");
    }
    return CLSPMV_OK;
}

int DevBuf::alloc(size_t bytes) {
    size = bytes;
    hipError_t e = hipMalloc(&ptr, bytes ? bytes : 4);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_status(e, "hipMalloc");
    }
    return CLSPMV_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    size = 0;
}

}  // namespace clspmv

struct clspmv_part_s {
    clspmv_ctx_t ctx = nullptr;
    clspmv_part_info info{};
    std::vector<DevBuf> bufs;
    // one of these is live, per info.format
    DevEll ell{};
    DevSell sell{};
    DevCsr csr{};
    DevCsrAdaptive csra{};
    DevCsrSliced csrs{};
    long long carry_bytes = 0;
    double codec_bound = 0.0;  // x-tiled encoder's bound, copied into info by finish_xtile
    long long xt_tiles = 0, xt_staged_bytes = 0;  // x-tiled: tiles and x bytes staged per SpMV
    int csr_lanes = 0;
    DevCoo coo{};
    DevDia dia{};
    DevBdia bdia{};
    DevBell bell{};
    DevSbell sbell{};
    DevBcsr bcsr{};
    DevBsr bsr{};
    DevXtile xt{};
    DevDiaMixed diam{};
    DevBsrMixed bsrm{};
    bool mixed = false;  // a mixed-precision part (diam, bsrm, or xt with a codec)

    void* add(clspmv_ctx_t c, const void* host, size_t bytes, int* rc) {
        bufs.emplace_back();
        *rc = bufs.back().upload(c, host, bytes);
        return bufs.back().ptr;
    }
    void* scratch(size_t bytes, int* rc) {
        bufs.emplace_back();
        *rc = bufs.back().alloc(bytes);
        return bufs.back().ptr;
    }
    ~clspmv_part_s() {
        for (auto& b : bufs) b.release();
    }
};

namespace {

void fill_info(clspmv_part_t p, int fmt, int method, int rows, int cols, long long nnz, const char* kname,
               int launches) {
    p->info.format = fmt;
    p->info.method = method;
    p->info.rows = rows;
    p->info.cols = cols;
    p->info.nnz = nnz;
    long long stored = 0;
    for (auto& b : p->bufs) stored += (long long)b.size;
    p->info.stored_bytes = stored;
    p->info.alg_bytes = stored + 4LL * cols + 8LL * rows;
    p->info.launches = launches;
    std::snprintf(p->info.kernel, sizeof(p->info.kernel), "%s", kname);
}

int pow2_floor(int v) {
    int p = 1;
    while (p * 2 <= v) p *= 2;
    return p;
}

bool cols_ok(const int* c, size_t n, int lo, int hi) {
    for (size_t i = 0; i < n; ++i)
        if (c[i] < lo || c[i] >= hi) return false;
    return true;
}

template <class F>
int make_part(clspmv_ctx_t ctx, clspmv_part_t* out, F build) {
    if (!ctx || !out) return CLSPMV_E_INVALID;
    clspmv_part_t p = new (std::nothrow) clspmv_part_s();
    if (!p) return CLSPMV_E_NOMEM;
    p->ctx = ctx;
    int rc = build(p);
    if (rc) {
        delete p;
        *out = nullptr;
        return rc;
    }
    *out = p;
    return CLSPMV_OK;
}

constexpr int kMethodMixed = 16;  // part method = 16 + codec (+ 8 with int16 delta columns)

static_assert(kCodecBf16x2 == CLSPMV_VAL_BF16X2 && kCodecFp16Row == CLSPMV_VAL_FP16_ROW &&
                  kCodecFp8Slice == CLSPMV_VAL_FP8_SLICE && kCodecI16Row == CLSPMV_VAL_I16_ROW,
              "codec ids are the ABI's CLSPMV_VAL_* values");
bool codec_ok(int codec) {
    return codec == kCodecBf16x2 || codec == kCodecFp16Row || codec == kCodecFp8Slice || codec == kCodecI16Row;
}

// BF16 hi + lo split of v: (bf16(v) << 16) | bf16(v - bf16(v)).
uint32_t bf16x2_word(float v) {
    const uint16_t hi = f32_to_bf16(v);
    const uint16_t lo = f32_to_bf16(v - bf16_to_f32(hi));
    return ((uint32_t)hi << 16) | lo;
}

// FP8 scale: the largest magnitude maps to the largest finite E4M3 (448).
float fp8_scale(float amax) { return amax > 0.f ? amax / 448.0f : 1.0f; }
// FP16 row scale: a power of two (so v / s is exact) that puts the row's
// largest magnitude in [2^14, 2^15), well inside FP16's normal range.
float fp16_scale(float amax) { return amax > 0.f ? std::ldexp(1.0f, std::ilogb(amax) - 14) : 1.0f; }
// Row scale of the int16 codec: the row's largest magnitude maps to 32767.
float i16_scale(float amax) { return amax > 0.f ? amax / 32767.0f : 1.0f; }
float row_scale(int codec, float amax) { return codec == kCodecI16Row ? i16_scale(amax) : fp16_scale(amax); }
// The int16 codec holds finite values only: a row with an infinite or NaN
// value is refused rather than saturated.
int check_row_scales(int codec, const std::vector<float>& s) {
    if (codec != kCodecI16Row) return CLSPMV_OK;
    for (float v : s)
        if (!std::isfinite(v)) {
            set_error("the int16 value codec needs finite matrix values");
            return CLSPMV_E_INVALID;
        }
    return CLSPMV_OK;
}
// v / s for a row-scaled codec, as the stored 16 bits.
uint16_t row_encode(int codec, float v_over_s) {
    if (codec == kCodecFp16Row) return f32_to_f16(v_over_s);
    return (uint16_t)(int16_t)std::max(-32767.0f, std::min(32767.0f, std::nearbyint(v_over_s)));
}

// Host decoders (the device's, kernels/kcodec.hpp) and the relative error
// of one encoded value, for the codec bound of clspmv_part_info.
float bf16x2_decode(uint32_t w) { return bf16_to_f32((uint16_t)(w >> 16)) + bf16_to_f32((uint16_t)(w & 0xffffu)); }
float row_decode(int codec, uint16_t h) {
    if (codec == kCodecFp16Row) {
        _Float16 f;
        std::memcpy(&f, &h, 2);
        return (float)f;
    }
    return (float)(int16_t)h;
}
// |dec - v| / |v|; a value encoded as anything but itself when v == 0 (or
// a non-finite v) counts as 1.
double rel_err(float v, double dec) {
    if (v == 0.0f) return dec == 0.0 ? 0.0 : 1.0;
    if (!std::isfinite(v)) return dec == (double)v ? 0.0 : 1.0;
    return std::min(1.0, std::fabs(dec - (double)v) / std::fabs((double)v));
}

const char* codec_name(int codec) {
    return codec == kCodecBf16x2  ? "bf16x2"
           : codec == kCodecFp16Row ? "fp16row"
           : codec == kCodecI16Row  ? "i16row"
                                    : "fp8slice";
}

// LDS-windowed sliced BSR kernels (tunable: "bsr_lds_window", 0 / 1).
// 0: never; 1 (default): by the rules of bsr_windows; 2: whenever every
// window fits LDS (no size or width rule; for tests on small matrices).
int g_bsr_window = 1;
constexpr int kBsrWinMinGroups = 128;

// Column repeated by the padding slots of an empty block row of the sliced
// BSR layout: the first column of the slice's first nonempty block row
// (inside the x window of the row's workgroup), else 0.
std::vector<int> bsr_slice_fill(const clspmv_bsr_matrix* m, int brows) {
    const int ns = (brows + kBsrSlice - 1) / kBsrSlice;
    std::vector<int> fill((size_t)ns, 0);
    for (int s = 0; s < ns; ++s)
        for (int r = s * kBsrSlice; r < std::min(brows, s * kBsrSlice + kBsrSlice); ++r)
            if (m->bsr_row_ptr[r] < m->bsr_row_ptr[r + 1]) {
                fill[(size_t)s] = m->bsr_col_id[m->bsr_row_ptr[r]];
                break;
            }
    return fill;
}

// x windows of the LDS-windowed sliced BSR kernels: for each workgroup of
// kBsrWinRows block rows, {smallest slot column, largest slot column + bw -
// smallest}.  Empty (the plain kernel is used) when the option is off, x
// has fewer than 4 columns (or so many that the staging's int column
// arithmetic, window start + 4 kBsrWinFloats, could overflow), some window
// holds more than kBsrWinFloats, or
// the windows are narrow: a mean under 2 kBsrWinRows bw columns (twice the
// rows' own width), or there are fewer than kBsrWinMinGroups windows (one
// 1024-lane workgroup per window: 4096 block rows would run on 4 of the
// 256 CUs).  There the plain kernel's gathers hit few cache lines
// per wave and its higher occupancy wins (1M 3x3 block rows, 9 blocks each
// within +-b block columns: plain 71-74 us for b <= 200, 85 us at b = 1000,
// 98 us at b = 5000; windowed 81 us throughout).
std::vector<int2> bsr_windows(const std::vector<int>& sp, const std::vector<int>& scol, int brows, int bw,
                              int cols) {
    if (!g_bsr_window || brows == 0 || cols < 4 || cols > INT32_MAX - 4 * kBsrWinFloats) return {};
    const int ns = (int)sp.size() - 1, per = kBsrWinRows / kBsrSlice;
    const int nwg = (brows + kBsrWinRows - 1) / kBsrWinRows;
    const bool rules = g_bsr_window == 1;
    if (rules && nwg < kBsrWinMinGroups) return {};
    std::vector<int2> w((size_t)nwg);
    long long total = 0;
    for (int g = 0; g < nwg; ++g) {
        int lo = INT32_MAX, hi = -1;
        for (int s = g * per; s < std::min(ns, g * per + per); ++s)
            for (int i = sp[(size_t)s]; i < sp[(size_t)s + 1]; ++i) {
                lo = std::min(lo, scol[(size_t)i]);
                hi = std::max(hi, scol[(size_t)i]);
            }
        if (hi < 0) {
            w[(size_t)g] = make_int2(0, 0);
            continue;
        }
        const long long len = (long long)hi + bw - lo;
        if (len > kBsrWinFloats) return {};
        w[(size_t)g] = make_int2(lo, (int)len);
        total += len;
    }
    if (rules && total < 2LL * kBsrWinRows * bw * nwg) return {};
    return w;
}

int build_csr_sliced(clspmv_ctx_t ctx, clspmv_part_t p, const int* rp, const int* ci, const float* cv, int rows,
                     int cols);
int build_xtile(clspmv_ctx_t ctx, clspmv_part_t p, const int* rp, const int* ci, const float* cv, int rows, int cols,
                int codec = 0);
int build_bsr_coded(clspmv_ctx_t ctx, clspmv_part_t p, const clspmv_bsr_matrix* m, int codec, int delta_cols);

// Row-list view (CSR arrays) of a stored format, for the layouts that are
// rebuilt from rows (column-sliced, x-tiled).  Every stored slot is kept,
// padding included, in slot order (= column order for the reference's
// conversions).
struct RowLists {
    std::vector<int> rp, ci;
    std::vector<float> cv;
};

int ell_rows(const clspmv_ell_matrix* m, RowLists* o) {
    const int rows = m->matinfo.height, K = m->ell_num, H = m->ell_height_aligned;
    if ((long long)rows * K > INT32_MAX) return CLSPMV_E_TOO_LARGE;
    o->rp.resize((size_t)rows + 1);
    o->ci.resize((size_t)rows * K);
    o->cv.resize((size_t)rows * K);
    for (int r = 0; r <= rows; ++r) o->rp[(size_t)r] = r * K;
#pragma omp parallel for schedule(static) num_threads(omp_threads())
    for (int r = 0; r < rows; ++r)
        for (int k = 0; k < K; ++k) {
            o->ci[(size_t)r * K + k] = m->ell_col_id[(size_t)k * H + r];
            o->cv[(size_t)r * K + k] = m->ell_data[(size_t)k * H + r];
        }
    return CLSPMV_OK;
}

// Method number of the LDS x-tiled layout, accepted by the ELL, SELL, CSR
// and COO parts.
constexpr int kMethodXtile = 4;

// Info of a part in the column-sliced layout (method 2); the long-row
// carry buffer is scratch, not format storage.
int finish_sliced(clspmv_part_t p, int fmt, int rows, int cols, long long nnz, const char* what) {
    char name[48];
    std::snprintf(name, sizeof(name), "%s_sliced_S%d", what, p->csrs.nslices);
    fill_info(p, fmt, 2, rows, cols, nnz, name, p->csrs.nslices + (p->csrs.n_long ? 1 : 0));
    p->info.stored_bytes -= p->carry_bytes;
    p->info.alg_bytes -= p->carry_bytes;
    return CLSPMV_OK;
}

int finish_xtile(clspmv_part_t p, int fmt, int rows, int cols, long long nnz) {
    char name[48];
    const int c = p->xt.codec;
    std::snprintf(name, sizeof(name), "xtile%s%s_B%d_G%d", c ? "_" : "", c ? codec_name(c) : "",
                  p->xt.nblocks / p->xt.groups, p->xt.groups);
    fill_info(p, fmt, c ? kMethodMixed + c : kMethodXtile, rows, cols, nnz, name, p->xt.groups > 1 ? 2 : 1);
    p->info.codec_bound = p->codec_bound;
    // the partial-sum buffer is scratch, not format storage
    p->info.stored_bytes -= p->carry_bytes;
    p->info.alg_bytes -= p->carry_bytes;
    return CLSPMV_OK;
}

}  // namespace

// ---------------------------------------------------------------- allocation-free COO
// clspmv_coo_spmv_noalloc keeps the COO kernels' chunk carries (two per
// 512-entry chunk: an int row and an FP64 sum) in this array instead of a
// device allocation: at most kNoAllocChunks chunks (nnz <= 167,772,160),
// 7.5 MiB held by every device the library runs on until the process ends.
// Host code uses the per-device address from hipGetSymbolAddress, never the
// array's name (which in host code is only a host-side stand-in).
namespace {
constexpr int kNoAllocChunks = 327680;
constexpr size_t kNoAllocRowBytes = (size_t)2 * kNoAllocChunks * sizeof(int);
constexpr size_t kNoAllocBytes = kNoAllocRowBytes + (size_t)2 * kNoAllocChunks * sizeof(double);
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};  // device address of noalloc_scratch, per device
}  // namespace

static __device__ __align__(256) unsigned char noalloc_scratch[kNoAllocBytes];

// Once per device, with that device current (called by clspmv_init).
static int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device %d beyond the %d the library supports", device, kMaxDevices);
        return CLSPMV_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return CLSPMV_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch));
    if (e != hipSuccess) return hip_status(e, "hipGetSymbolAddress(noalloc_scratch)");
    if (!ptr) {
        set_error("hipGetSymbolAddress(noalloc_scratch) returned no address");
        return CLSPMV_E_HIP;
    }
    g_noalloc_ptr[device] = ptr;
    return CLSPMV_OK;
}

extern "C" {

int clspmv_init(int device, clspmv_ctx_t* ctx) {
    if (!ctx) return CLSPMV_E_INVALID;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error("no HIP device visible");
        return CLSPMV_E_NODEVICE;
    }
    if (device < 0 || device >= n) return CLSPMV_E_INVALID;
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_status(e, "hipSetDevice");
    if (int rc = noalloc_setup(device)) return rc;
    auto* c = new (std::nothrow) clspmv_ctx_s();
    if (!c) return CLSPMV_E_NOMEM;
    c->device = device;
    // Default priority stream (never a prioritised one).
    if ((e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking)) != hipSuccess) {
        delete c;
        return hip_status(e, "hipStreamCreate");
    }
    if ((e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess) {
        (void)hipStreamDestroy(c->own);
        delete c;
        return hip_status(e, "hipEventCreate");
    }
    *ctx = c;
    return CLSPMV_OK;
}

int clspmv_finalize(clspmv_ctx_t ctx) {
    if (!ctx) return CLSPMV_E_INVALID;
    (void)hipStreamSynchronize(ctx->stream());
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipStreamDestroy(ctx->own);
    delete ctx;
    return CLSPMV_OK;
}

int clspmv_set_stream(clspmv_ctx_t ctx, void* stream) {
    if (!ctx) return CLSPMV_E_INVALID;
    ctx->ext = reinterpret_cast<hipStream_t>(stream);
    return CLSPMV_OK;
}

int clspmv_synchronize(clspmv_ctx_t ctx) {
    if (!ctx) return CLSPMV_E_INVALID;
    return hip_status(hipStreamSynchronize(ctx->stream()), "hipStreamSynchronize");
}

int clspmv_device_alloc(clspmv_ctx_t ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return CLSPMV_E_INVALID;
    return hip_status(hipMalloc(dptr, bytes ? bytes : 4), "hipMalloc");
}
int clspmv_device_free(clspmv_ctx_t ctx, void* dptr) {
    if (!ctx) return CLSPMV_E_INVALID;
    return hip_status(hipFree(dptr), "hipFree");
}
int clspmv_memcpy_h2d(clspmv_ctx_t ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return CLSPMV_E_INVALID;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream());
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream());
    return hip_status(e, "hipMemcpy H2D");
}
int clspmv_memcpy_d2h(clspmv_ctx_t ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return CLSPMV_E_INVALID;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream());
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream());
    return hip_status(e, "hipMemcpy D2H");
}
int clspmv_memcpy_d2d(clspmv_ctx_t ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return CLSPMV_E_INVALID;
    return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream()), "hipMemcpy D2D");
}

int clspmv_event_elapsed_begin(clspmv_ctx_t ctx) {
    if (!ctx) return CLSPMV_E_INVALID;
    return hip_status(hipEventRecord(ctx->ev0, ctx->stream()), "hipEventRecord");
}
int clspmv_event_elapsed_end(clspmv_ctx_t ctx, float* ms) {
    if (!ctx || !ms) return CLSPMV_E_INVALID;
    hipError_t e = hipEventRecord(ctx->ev1, ctx->stream());
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, ctx->ev0, ctx->ev1);
    return hip_status(e, "hipEvent timing");
}

// ---------------------------------------------------------------- parts

int clspmv_part_ell(clspmv_ctx_t ctx, const clspmv_ell_matrix* m, int method, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        const int rows = m->matinfo.height, cols = m->matinfo.width, H = m->ell_height_aligned, K = m->ell_num;
        if (rows < 0 || H < rows || K < 0 || (H % 4) != 0) return CLSPMV_E_INVALID;
        // Every stored slot is checked, including the padding rows up to H:
        // the four-rows-per-lane kernel reads them.
        if (rows > 0 && K > 0 && !cols_ok(m->ell_col_id, (size_t)H * K, 0, cols)) {
            set_error("ELL column index out of range");
            return CLSPMV_E_INVALID;
        }
        int rc = 0;
        size_t n = (size_t)H * K;
        if (method == kMethodXtile) {
            RowLists rl;
            if ((rc = ell_rows(m, &rl))) return rc;
            if ((rc = build_xtile(ctx, p, rl.rp.data(), rl.ci.data(), rl.cv.data(), rows, cols))) return rc;
            return finish_xtile(p, CLSPMV_FMT_ELL, rows, cols, m->matinfo.nnz);
        }
        if (method == 2) {
            // Column-sliced device layout built from the ELL slots.
            RowLists rl;
            if ((rc = ell_rows(m, &rl))) return rc;
            rc = build_csr_sliced(ctx, p, rl.rp.data(), rl.ci.data(), rl.cv.data(), rows, cols);
            if (rc) return rc;
            char name[32];
            std::snprintf(name, sizeof(name), "ell_sliced_S%d", p->csrs.nslices);
            fill_info(p, CLSPMV_FMT_ELL, 2, rows, cols, m->matinfo.nnz, name, p->csrs.nslices + (p->csrs.n_long ? 1 : 0));
            p->info.stored_bytes -= p->carry_bytes;
            p->info.alg_bytes -= p->carry_bytes;
            return CLSPMV_OK;
        }
        p->ell.col = (const int*)p->add(ctx, m->ell_col_id, n * 4, &rc);
        if (rc) return rc;
        p->ell.val = (const float*)p->add(ctx, m->ell_data, n * 4, &rc);
        if (rc) return rc;
        p->ell.H = H;
        p->ell.ell_num = K;
        p->ell.rows = rows;
        int meth = (method == 1 || method == 3) ? 1 : 0;  // 3: texture variant of 1
        fill_info(p, CLSPMV_FMT_ELL, meth, rows, cols, m->matinfo.nnz, meth ? "ell_row4" : "ell_row", 1);
        return CLSPMV_OK;
    });
}

int clspmv_part_sell(clspmv_ctx_t ctx, const clspmv_sell_matrix* m, int method, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        const int rows = m->matinfo.height, cols = m->matinfo.width;
        const int S = (int)m->sell_slice_height, ns = m->sell_slice_num;
        if (S <= 0 || ns != (rows + S - 1) / S) return CLSPMV_E_INVALID;
        if (m->sell_slice_ptr[0] != 0) return CLSPMV_E_INVALID;
        for (int s = 0; s < ns; ++s) {
            int a = m->sell_slice_ptr[s], b = m->sell_slice_ptr[s + 1];
            if (b < a || (b - a) % S) return CLSPMV_E_INVALID;
            int w = (b - a) / S;
            for (int r = s * S; r < std::min(rows, s * S + S); ++r)
                for (int j = 0; j < w; ++j) {
                    int c = m->sell_col_id[(size_t)a + (size_t)j * S + (r - s * S)];
                    if (c < 0 || c >= cols) {
                        set_error("SELL column index out of range");
                        return CLSPMV_E_INVALID;
                    }
                }
        }
        int rc = 0;
        if (method == kMethodXtile || method == 2) {
            // rows of the slices, padding slots included
            RowLists rl;
            rl.rp.assign((size_t)rows + 1, 0);
            for (int s = 0; s < ns; ++s) {
                const int w = (m->sell_slice_ptr[s + 1] - m->sell_slice_ptr[s]) / S;
                for (int r = s * S; r < std::min(rows, s * S + S); ++r) rl.rp[(size_t)r + 1] = rl.rp[(size_t)r] + w;
            }
            rl.ci.resize((size_t)rl.rp[(size_t)rows]);
            rl.cv.resize((size_t)rl.rp[(size_t)rows]);
            for (int s = 0; s < ns; ++s) {
                const int a = m->sell_slice_ptr[s], w = (m->sell_slice_ptr[s + 1] - a) / S;
                for (int r = s * S; r < std::min(rows, s * S + S); ++r)
                    for (int j = 0; j < w; ++j) {
                        const size_t src = (size_t)a + (size_t)j * S + (size_t)(r - s * S);
                        rl.ci[(size_t)rl.rp[(size_t)r] + j] = m->sell_col_id[src];
                        rl.cv[(size_t)rl.rp[(size_t)r] + j] = m->sell_data[src];
                    }
            }
            if (method == 2) {
                rc = build_csr_sliced(ctx, p, rl.rp.data(), rl.ci.data(), rl.cv.data(), rows, cols);
                if (rc) return rc;
                return finish_sliced(p, CLSPMV_FMT_SELL, rows, cols, m->matinfo.nnz, "sell");
            }
            if ((rc = build_xtile(ctx, p, rl.rp.data(), rl.ci.data(), rl.cv.data(), rows, cols))) return rc;
            return finish_xtile(p, CLSPMV_FMT_SELL, rows, cols, m->matinfo.nnz);
        }
        size_t total = (size_t)m->sell_slice_ptr[ns];
        p->sell.slice_ptr = (const int*)p->add(ctx, m->sell_slice_ptr, ((size_t)ns + 1) * 4, &rc);
        if (rc) return rc;
        p->sell.col = (const int*)p->add(ctx, m->sell_col_id, total * 4, &rc);
        if (rc) return rc;
        p->sell.val = (const float*)p->add(ctx, m->sell_data, total * 4, &rc);
        if (rc) return rc;
        p->sell.S = S;
        p->sell.rows = rows;
        fill_info(p, CLSPMV_FMT_SELL, method & 1, rows, cols, m->matinfo.nnz, "sell_row", 1);
        return CLSPMV_OK;
    });
}

static int validate_csr(const clspmv_csr_matrix* m) {
    const int rows = m->matinfo.height, cols = m->matinfo.width;
    if (rows < 0 || m->csr_row_ptr[0] != 0) return CLSPMV_E_INVALID;
    for (int r = 0; r < rows; ++r)
        if (m->csr_row_ptr[r + 1] < m->csr_row_ptr[r]) return CLSPMV_E_INVALID;
    if (!cols_ok(m->csr_col_id, (size_t)m->csr_row_ptr[rows], 0, cols)) {
        set_error("CSR column index out of range");
        return CLSPMV_E_INVALID;
    }
    return CLSPMV_OK;
}

static int upload_csr(clspmv_ctx_t ctx, clspmv_part_t p, const clspmv_csr_matrix* m) {
    int rc = validate_csr(m);
    if (rc) return rc;
    const int rows = m->matinfo.height;
    const size_t nnz = (size_t)m->csr_row_ptr[rows];
    p->csr.row_ptr = (const int*)p->add(ctx, m->csr_row_ptr, ((size_t)rows + 1) * 4, &rc);
    if (rc) return rc;
    p->csr.col = (const int*)p->add(ctx, m->csr_col_id, nnz * 4, &rc);
    if (rc) return rc;
    p->csr.val = (const float*)p->add(ctx, m->csr_data, nnz * 4, &rc);
    if (rc) return rc;
    p->csr.rows = rows;
    return CLSPMV_OK;
}

}  // extern "C"

namespace {

// Column slices whose x range is at most ~2 MiB: small enough to stay in an
// XCD's 4 MiB L2 next to the streaming entries.
long long g_slice_x_bytes = 2ll << 20;  // tunable: clspmv_set_option("col_slice_bytes", v)

// Builds the column-sliced device layout (DevCsrSliced) from CSR arrays whose
// rows are sorted by column.  Blocks per slice follow the adaptive planner;
// LONG chunks get carry slots grouped by row (a row long in several slices
// has consecutive slots), so the finalize kernel has one writer per row.
int build_csr_sliced(clspmv_ctx_t ctx, clspmv_part_t p, const int* rp, const int* ci, const float* cv, int rows,
                     int cols) {
    const long long nnz = rows ? rp[rows] : 0;
    const long long sxb = std::max(256ll, g_slice_x_bytes);
    int ns = (int)std::min<long long>(kMaxColSlices, std::max<long long>(1, (4ll * cols + sxb - 1) / sxb));
    const int sw = std::max(1, (cols + ns - 1) / ns);  // columns per slice
    ns = std::max(1, (cols + sw - 1) / sw);
    std::vector<int> cnt((size_t)ns * rows, 0);
#pragma omp parallel for schedule(static) num_threads(omp_threads())
    for (int r = 0; r < rows; ++r)
        for (int k = rp[r]; k < rp[r + 1]; ++k) cnt[(size_t)(ci[k] / sw) * rows + r]++;
    // slice-major start of every (slice, row) segment
    std::vector<long long> start((size_t)ns * rows + 1, 0);
    {
        long long acc = 0;
        for (size_t i = 0; i < (size_t)ns * rows; ++i) {
            start[i] = acc;
            acc += cnt[i];
        }
        start[(size_t)ns * rows] = acc;
    }
    if (nnz > INT32_MAX) return CLSPMV_E_TOO_LARGE;
    std::vector<int> scol((size_t)nnz);
    std::vector<float> sval((size_t)nnz);
#pragma omp parallel for schedule(static) num_threads(omp_threads())
    for (int r = 0; r < rows; ++r) {
        int s_prev = -1;
        long long pos = 0;
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
            const int s = ci[k] / sw;
            if (s != s_prev) {
                pos = start[(size_t)s * rows + r];
                s_prev = s;
            }
            scol[(size_t)pos] = ci[k];
            sval[(size_t)pos] = cv[k];
            ++pos;
        }
    }
    std::vector<unsigned char> cb((size_t)ns * rows);
    for (size_t i = 0; i < cb.size(); ++i) cb[i] = (unsigned char)std::min(255, cnt[i]);
    // blocks
    std::vector<std::vector<int4>> sb((size_t)ns);
    std::vector<std::pair<int, int>> longs;  // (row, slice)
    for (int s = 0; s < ns; ++s) {
        const int* c = &cnt[(size_t)s * rows];
        const long long* st = &start[(size_t)s * rows];
        int r = 0;
        while (r < rows) {
            const int len = c[r];
            if (len > kCsrStreamMaxRowLen) {
                if (len <= kCsrLongChunk)
                    sb[s].push_back(make_int4(r, r + 1, (int)st[r], (int)(st[r] + len)));
                else
                    longs.push_back({r, s});
                ++r;
                continue;
            }
            const int r0 = r;
            int n = 0;
            while (r < rows && r - r0 < 64) {
                const int l = c[r];
                if (l > kCsrStreamMaxRowLen || n + l > kCsrStreamMaxNnz) break;
                n += l;
                ++r;
            }
            sb[s].push_back(make_int4(r0, r, (int)st[r0], (int)(st[r0] + n)));
        }
    }
    std::sort(longs.begin(), longs.end());
    std::vector<int> long_rows, long_first{0};
    int slots = 0;
    for (size_t i = 0; i < longs.size(); ++i) {
        const int r = longs[i].first, s = longs[i].second;
        const long long a = start[(size_t)s * rows + r], e = a + cnt[(size_t)s * rows + r];
        for (long long n0 = a; n0 < e; n0 += kCsrLongChunk) {
            const long long n1 = std::min(e, n0 + kCsrLongChunk);
            sb[s].push_back(make_int4(r, -(slots + 1), (int)n0, (int)n1));
            ++slots;
        }
        if (i + 1 == longs.size() || longs[i + 1].first != r) {
            long_rows.push_back(r);
            long_first.push_back(slots);
        }
    }
    std::vector<int4> blocks;
    DevCsrSliced d{};
    for (int s = 0; s < ns; ++s) {
        d.slice_block_begin[s] = (int)blocks.size();
        blocks.insert(blocks.end(), sb[s].begin(), sb[s].end());
    }
    d.slice_block_begin[ns] = (int)blocks.size();
    int rc = 0;
    d.col = (const int*)p->add(ctx, scol.data(), scol.size() * 4, &rc);
    if (rc) return rc;
    d.val = (const float*)p->add(ctx, sval.data(), sval.size() * 4, &rc);
    if (rc) return rc;
    d.counts = (const unsigned char*)p->add(ctx, cb.data(), cb.size(), &rc);
    if (rc) return rc;
    d.blocks = (const int4*)p->add(ctx, blocks.data(), blocks.size() * sizeof(int4), &rc);
    if (rc) return rc;
    d.long_rows = (const int*)p->add(ctx, long_rows.data(), long_rows.size() * 4, &rc);
    if (rc) return rc;
    d.long_first = (const int*)p->add(ctx, long_first.data(), long_first.size() * 4, &rc);
    if (rc) return rc;
    d.carry = (double*)p->scratch((size_t)std::max(1, slots) * 8, &rc);
    if (rc) return rc;
    d.nslices = ns;
    d.rows = rows;
    d.n_long = (int)long_rows.size();
    p->csrs = d;
    p->carry_bytes = (long long)std::max(1, slots) * 8;
    return CLSPMV_OK;
}

int g_xt_blocks = 256;  // target workgroups: one per CU (tunable: "xtile_blocks")

// A row's entries in one column slice (a segment of L entries) are placed
// by class: 0 (L <= 2) first, in row order, then 1 (3 <= L <= kXtAlignMax),
// then 2 (longer), so that chunks of short rows take the kernel's fast path
// and the others need few scan steps.  Where a segment starts, given its
// slice's next free position p (slices start on chunk boundaries, so p mod
// kXtChunk is the offset in the chunk): segments of 2 to kXtAlignMax
// entries start at an even offset (a row of 2 entries in one lane), and
// class 1 does not straddle a chunk boundary (its sum needs no carry).  The
// positions skipped are padding.  A slice is placed without this padding
// (`aligned` false) when the padding would cost its tiles another round of
// chunks (xt_rounds).
constexpr int kXtAlignMax = 16;
inline int xt_class(long long L) { return L <= 2 ? 0 : (L <= kXtAlignMax ? 1 : 2); }
inline long long xt_rounds(long long slots) { return ((slots + kXtChunk - 1) / kXtChunk + kXtWaves - 1) / kXtWaves; }
inline long long xt_place(long long p, long long L, bool aligned) {
    const int cls = xt_class(L);
    if (!aligned || cls == 2 || L == 1) return p;
    p += p & 1;
    const long long off = p % kXtChunk;
    if (cls == 1 && off + L > kXtChunk) p += kXtChunk - off;
    return p;
}

// Visits the rows r0 .. r1 - 1 in order, three times (segment classes 0, 1,
// 2); in pass c, for each row, first calls place(row, slice, L) for every
// column slice in which the row has a segment of class c (L entries), then
// entry(row, k, slice) for each entry of those segments.  rowcnt (ns zeros)
// and touched are scratch.
template <class Place, class Entry>
void xt_walk_rows(const int* rp, const int* ci, int r0, int r1, std::vector<int>& rowcnt, std::vector<int>& touched,
                  Place place, Entry entry) {
    for (int cls = 0; cls < 3; ++cls)
        for (int r = r0; r < r1; ++r) {
            for (long long k = rp[r]; k < rp[r + 1]; ++k) {
                const int sl = ci[k] / kXtCols;
                if (rowcnt[(size_t)sl]++ == 0) touched.push_back(sl);
            }
            for (int sl : touched)
                if (xt_class(rowcnt[(size_t)sl]) == cls) place(r, sl, (long long)rowcnt[(size_t)sl]);
            for (long long k = rp[r]; k < rp[r + 1]; ++k) {
                const int sl = ci[k] / kXtCols;
                if (xt_class(rowcnt[(size_t)sl]) == cls) entry(r, k, sl);
            }
            for (int sl : touched) rowcnt[(size_t)sl] = 0;
            touched.clear();
        }
}

// Column groups per row block at most (tunable: "xtile_max_groups", 1..8).
int g_xt_max_groups = kXtMaxGroups;
// Issue priority of the x-tiled kernel's waves on scan chunks: -1 by the
// layout's share of fast chunks, else 0 or 1 (tunable: "xtile_scan_prio").
int g_xt_scan_prio = -1;

// Builds the LDS x-tiled layout (DevXtile) from row lists.  Every row block
// stages the x it touches once, so the x traffic is (row blocks) x (x
// bytes): the blocks are as tall as LDS allows (kXtRows), at least W / 8 of
// them (W = target workgroups), balanced by entry count; each block is then
// split into G = W / blocks (<= 8) column groups so that the launch still
// has about W workgroups.  When balancing by entries would need more
// blocks than that minimum and so lower G, the rows are cut evenly instead.
// Within a block the entries are bucketed by column slice, each row's
// entries of a slice contiguous (in column order when the row's are), the
// rows ordered by segment class and then by row (xt_place).  Each chunk is
// flagged for the kernel's fast path when no row crosses one of its lane
// boundaries, with the carry bit when its first row continues the previous
// chunk of the tile (a tile with a carry is flagged too), and otherwise
// carries the plan of its segmented scan.
// `codec` (0 or a kCodec* value) encodes the values: FP16 divided by a
// power-of-two scale per row, int16 fixed point with one scale per row, FP8
// with one scale per 128-entry chunk.
// The x-tiled layout's host arrays (everything build_xtile uploads, before
// the FP32 interleave and the value codecs' encoding; values divided by
// the row scale for row-scaled codecs).
struct XtHostLayout {
    std::vector<int> brow, blk_tile, tile_slice, tile_chunk;
    std::vector<unsigned> idx;
    std::vector<float> val, rscale;
    int nb = 0, G = 1;
    long long fast_chunks = 0;  // chunks with the fast bit
};

int xt_host_layout(const int* rp, const int* ci, const float* cv, int rows, int cols, int codec, XtHostLayout* out) {
    const long long nnz = rows ? rp[rows] : 0;
    if (cols < 4) {
        set_error("the x-tiled layout needs at least 4 columns");
        return CLSPMV_E_UNSUPPORTED;
    }
    const int ns = std::max(1, (cols + kXtCols - 1) / kXtCols);
    const int W = std::max(1, g_xt_blocks);
    const int max_g = std::max(1, std::min({g_xt_max_groups, kXtMaxGroups, ns}));
    const long long nb_rows = std::max(1, (rows + kXtRows - 1) / kXtRows);
    const long long nb_target = std::max(nb_rows, (long long)(W + max_g - 1) / max_g);
    std::vector<int> brow{0};
    const long long target = std::max<long long>(1, (nnz + nb_target - 1) / nb_target);
    for (int r = 0; r < rows;) {
        const int rb = r;
        long long n = 0;
        do {
            n += rp[r + 1] - rp[r];
            ++r;
        } while (r < rows && r - rb < kXtRows && n < target);
        brow.push_back(r);
    }
    auto groups_for = [&](long long nb) { return (int)std::max(1LL, std::min((long long)max_g, W / nb)); };
    if ((long long)brow.size() - 1 > nb_target && groups_for((long long)brow.size() - 1) < groups_for(nb_target)) {
        // even rows: nb_target blocks of at most kXtRows rows
        brow.assign(1, 0);
        const long long per = (rows + nb_target - 1) / nb_target;
        for (long long r = per; r < rows; r += per) brow.push_back((int)r);
        brow.push_back(rows);
    }
    const int nb = (int)brow.size() - 1;
    // per block: (slice, entries) of its nonempty slices
    struct XtSliceFill {
        int slice;
        long long slots;  // entries + padding
        bool aligned;     // placed with the alignment padding
    };
    std::vector<std::vector<XtSliceFill>> bt((size_t)nb);
#pragma omp parallel num_threads(omp_threads())
    {
        std::vector<long long> cnt((size_t)ns), plain((size_t)ns);
        std::vector<int> rowcnt((size_t)ns, 0), touched;
#pragma omp for schedule(dynamic, 4)
        for (int b = 0; b < nb; ++b) {
            std::fill(cnt.begin(), cnt.end(), 0);
            std::fill(plain.begin(), plain.end(), 0);
            // slots per slice, with and without the alignment padding
            xt_walk_rows(
                rp, ci, brow[b], brow[b + 1], rowcnt, touched,
                [&](int, int sl, long long L) {
                    cnt[(size_t)sl] = xt_place(cnt[(size_t)sl], L, true) + L;
                    plain[(size_t)sl] += L;
                },
                [](int, long long, int) {});
            for (int sl = 0; sl < ns; ++sl) {
                if (!plain[(size_t)sl]) continue;
                const bool aligned = xt_rounds(cnt[(size_t)sl]) <= xt_rounds(plain[(size_t)sl]);
                bt[(size_t)b].push_back({sl, aligned ? cnt[(size_t)sl] : plain[(size_t)sl], aligned});
            }
        }
    }
    // Column groups: G contiguous ranges of slices; workgroup
    // b = block * G + group.
    const int G = rows ? groups_for(nb) : 1;
    const int spg = (ns + G - 1) / G;
    // A slice's chunks are cut into tiles of at most kXtTileChunks; the
    // later pieces are marked as continuing the slice (x stays in LDS).
    std::vector<int> blk_tile{0}, tile_slice, tile_chunk{0};
    for (int b = 0; b < nb; ++b) {
        for (int g = 0; g < G; ++g) {
            for (auto& t : bt[(size_t)b]) {
                if (t.slice / spg != g) continue;
                const long long nc = (t.slots + kXtChunk - 1) / kXtChunk;
                if (((long long)tile_chunk.back() + nc) * kXtChunk > INT32_MAX) return CLSPMV_E_TOO_LARGE;
                for (long long k = 0; k < nc; k += kXtTileChunks) {
                    tile_slice.push_back((t.slice << 2) | (k ? 1 : 0));
                    tile_chunk.push_back(tile_chunk.back() + (int)std::min<long long>(kXtTileChunks, nc - k));
                }
            }
            blk_tile.push_back((int)tile_slice.size());
        }
    }
    if ((long long)nb * G > INT32_MAX / 2) return CLSPMV_E_TOO_LARGE;
    const size_t total = (size_t)tile_chunk.back() * kXtChunk;
    std::vector<unsigned> idx(std::max<size_t>(total, 1), kXtPadRow << 16);
    std::vector<float> val(std::max<size_t>(total, 1), 0.f);
    std::vector<float> rscale;
    if (codec_row_scaled(codec)) {
        rscale.assign((size_t)std::max(rows, 1), 1.0f);
#pragma omp parallel for schedule(static) num_threads(omp_threads())
        for (int r = 0; r < rows; ++r) {
            float a = 0.f;
            for (long long k = rp[r]; k < rp[r + 1]; ++k) a = std::max(a, std::fabs(cv[k]));
            rscale[(size_t)r] = row_scale(codec, a);
        }
        if (int e = check_row_scales(codec, rscale)) return e;
    }
#pragma omp parallel num_threads(omp_threads())
    {
        std::vector<long long> pos((size_t)ns, -1);
        std::vector<char> aligned((size_t)ns, 1);
        std::vector<int> rowcnt((size_t)ns, 0), touched;
#pragma omp for schedule(dynamic, 4)
        for (int b = 0; b < nb; ++b) {
            // entries of a slice run contiguously over its pieces
            for (int t = blk_tile[(size_t)b * G]; t < blk_tile[(size_t)(b + 1) * G]; ++t)
                if (!(tile_slice[(size_t)t] & 1))
                    pos[(size_t)(tile_slice[(size_t)t] >> 2)] = (long long)tile_chunk[(size_t)t] * kXtChunk;
            for (auto& f : bt[(size_t)b]) aligned[(size_t)f.slice] = f.aligned;
            xt_walk_rows(
                rp, ci, brow[b], brow[b + 1], rowcnt, touched,
                [&](int, int sl, long long L) { pos[(size_t)sl] = xt_place(pos[(size_t)sl], L, aligned[(size_t)sl]); },
                [&](int r, long long k, int sl) {
                    const long long q = pos[(size_t)sl]++;
                    idx[(size_t)q] = ((unsigned)(r - brow[b]) << 16) | (unsigned)(ci[k] - sl * kXtCols);
                    val[(size_t)q] = rscale.empty() ? cv[k] : cv[k] / rscale[(size_t)r];
                });
        }
    }
    // Chunk flags (k_xtile.hip), in the column halves of the chunk's first
    // two entries: the carry bit when the chunk's first row continues the
    // previous chunk of the tile (a padding slot never does); the fast bit
    // when there is no carry and no row crosses a lane boundary (entries
    // 2l - 1 and 2l in different rows, or padding); otherwise the scan plan
    // over the lanes' tail rows (entries 2l + 1, padding never matching):
    // in-row steps = ceil(log2 of the longest run of equal tail rows within
    // a 16-lane row), at least 1, and whether a run crosses lanes 15|16 or
    // 47|48 (bc15) and 31|32 (bc31).  A tile holding a carry gets bit 1 of
    // its slice word.
    long long fast_chunks = 0;
#pragma omp parallel for schedule(static) num_threads(omp_threads()) reduction(+ : fast_chunks)
    for (long long t = 0; t < (long long)tile_slice.size(); ++t) {
        bool carries = false;
        for (int c = tile_chunk[(size_t)t]; c < tile_chunk[(size_t)t + 1]; ++c) {
            const size_t q = (size_t)c * kXtChunk;
            auto row = [&](int j) { return idx[q + (size_t)j] >> 16; };
            const bool carry = c > tile_chunk[(size_t)t] && row(0) == (idx[q - 1] >> 16);
            bool fast = !carry;
            for (int l = 1; l < kXtChunk / 2 && fast; ++l) fast = row(2 * l) == kXtPadRow || row(2 * l) != row(2 * l - 1);
            if (carry) {
                idx[q] |= kXtCarryBit;
                carries = true;
            } else if (fast) {
                idx[q] |= kXtFastBit;
                ++fast_chunks;
                continue;
            }
            auto same_tail = [&](int l) { return row(2 * l + 1) != kXtPadRow && row(2 * l + 1) == row(2 * l - 1); };
            int run = 1, longest = 1;
            for (int l = 1; l < kXtChunk / 2; ++l) {
                run = (l % 16 != 0 && same_tail(l)) ? run + 1 : 1;
                longest = std::max(longest, run);
            }
            int steps = 1;
            while ((1 << steps) < longest) ++steps;
            idx[q + 1] |= (unsigned)(steps - 1) << kXtStepsShift;
            if (same_tail(16) || same_tail(48)) idx[q + 1] |= kXtBc15Bit;
            if (same_tail(32)) idx[q] |= kXtBc31Bit;
        }
        if (carries) tile_slice[(size_t)t] |= 2;
    }
    out->brow.swap(brow);
    out->blk_tile.swap(blk_tile);
    out->tile_slice.swap(tile_slice);
    out->tile_chunk.swap(tile_chunk);
    out->idx.swap(idx);
    out->val.swap(val);
    out->rscale.swap(rscale);
    out->nb = nb;
    out->G = G;
    out->fast_chunks = fast_chunks;
    return CLSPMV_OK;
}

int build_xtile(clspmv_ctx_t ctx, clspmv_part_t p, const int* rp, const int* ci, const float* cv, int rows,
                int cols, int codec) {
    XtHostLayout H;
    if (int e = xt_host_layout(rp, ci, cv, rows, cols, codec, &H)) return e;
    std::vector<int>&brow = H.brow, &blk_tile = H.blk_tile, &tile_slice = H.tile_slice, &tile_chunk = H.tile_chunk;
    std::vector<unsigned>& idx = H.idx;
    std::vector<float>&val = H.val, &rscale = H.rscale;
    const int nb = H.nb, G = H.G;
    int rc = 0;
    DevXtile d{};
    d.blk_row = (const int*)p->add(ctx, brow.data(), brow.size() * 4, &rc);
    if (rc) return rc;
    d.blk_tile = (const int*)p->add(ctx, blk_tile.data(), blk_tile.size() * 4, &rc);
    if (rc) return rc;
    d.tile_slice = (const int*)p->add(ctx, tile_slice.data(), std::max<size_t>(1, tile_slice.size()) * 4, &rc);
    if (rc) return rc;
    d.tile_chunk = (const int*)p->add(ctx, tile_chunk.data(), tile_chunk.size() * 4, &rc);
    if (rc) return rc;
    {
        // x bytes staged per SpMV: one slice for each tile that starts one
        long long starts = 0;
        for (size_t t = 0; t < tile_slice.size(); ++t) starts += !(tile_slice[t] & 1);
        p->xt_tiles = (long long)tile_slice.size();
        p->xt_staged_bytes = starts * (long long)kXtCols * 4;
    }
    // Per chunk and lane, the lane's two entries and its two values in one
    // record, one load per lane (k_xtile.hip: xt_fetch_e): FP32 and BF16
    // hi+lo 4 words, FP16 / int16 3 (the two 2-byte values in the third).
    // FP8 keeps entries and values apart (2 bytes of values per lane).
    const size_t nl = (size_t)tile_chunk.back() * 64;
    auto pack = [&](int words, auto value_word) {
        std::vector<uint32_t> ev(std::max<size_t>(nl, 1) * (size_t)words, 0);
#pragma omp parallel for schedule(static) num_threads(omp_threads())
        for (long long i = 0; i < (long long)nl; ++i) {
            uint32_t* o = ev.data() + (size_t)i * words;
            o[0] = idx[(size_t)i * 2];
            o[1] = idx[(size_t)i * 2 + 1];
            value_word((size_t)i, o + 2);
        }
        d.idx = (const unsigned*)p->add(ctx, ev.data(), ev.size() * 4, &rc);
        d.val = d.idx;
    };
    std::vector<float> vscale;
    double bound = 0.0;  // row-scaled codecs: the scale cancels in the ratio
    if (codec == 0) {
        pack(4, [&](size_t i, uint32_t* o) { std::memcpy(o, &val[i * 2], 8); });
    } else if (codec == kCodecBf16x2) {
        std::vector<uint32_t> w(val.size());
#pragma omp parallel for schedule(static) num_threads(omp_threads()) reduction(max : bound)
        for (long long i = 0; i < (long long)val.size(); ++i) {
            w[(size_t)i] = bf16x2_word(val[(size_t)i]);
            bound = std::max(bound, rel_err(val[(size_t)i], bf16x2_decode(w[(size_t)i])));
        }
        pack(4, [&](size_t i, uint32_t* o) { o[0] = w[i * 2], o[1] = w[i * 2 + 1]; });
    } else if (codec_row_scaled(codec)) {
        std::vector<uint16_t> h(val.size());
#pragma omp parallel for schedule(static) num_threads(omp_threads()) reduction(max : bound)
        for (long long i = 0; i < (long long)val.size(); ++i) {
            h[(size_t)i] = row_encode(codec, val[(size_t)i]);
            bound = std::max(bound, rel_err(val[(size_t)i], row_decode(codec, h[(size_t)i])));
        }
        pack(3, [&](size_t i, uint32_t* o) { o[0] = (uint32_t)h[i * 2] | ((uint32_t)h[i * 2 + 1] << 16); });
        vscale.swap(rscale);
    } else if (codec == kCodecFp8Slice) {
        d.idx = (const unsigned*)p->add(ctx, idx.data(), idx.size() * 4, &rc);
        if (rc) return rc;
        const long long nc = std::max(1, tile_chunk.back());
        std::vector<uint8_t> q((size_t)nc * kXtChunk, 0);
        vscale.assign((size_t)nc, 1.0f);
#pragma omp parallel for schedule(static) num_threads(omp_threads()) reduction(max : bound)
        for (long long c = 0; c < nc; ++c) {
            const float* v = val.data() + (size_t)c * kXtChunk;
            float a = 0.f;
            for (int i = 0; i < kXtChunk; ++i) a = std::max(a, std::fabs(v[i]));
            const float sc = fp8_scale(a);
            vscale[(size_t)c] = sc;
            for (int i = 0; i < kXtChunk; ++i) {
                const uint8_t e = f32_to_e4m3(v[i] / sc);
                q[(size_t)c * kXtChunk + (size_t)i] = e;
                bound = std::max(bound, rel_err(v[i], (double)sc * e4m3_to_f32(e)));
            }
        }
        d.val = p->add(ctx, q.data(), q.size(), &rc);
    }
    p->codec_bound = bound;
    if (rc) return rc;
    if (!vscale.empty()) {
        d.vscale = (const float*)p->add(ctx, vscale.data(), vscale.size() * 4, &rc);
        if (rc) return rc;
    }
    d.codec = codec;
    // when at least half the chunks take the fast path, the waves on scan
    // chunks are the ones each barrier waits for: they raise their issue
    // priority (tunable: "xtile_scan_prio" -1 this rule, 0 never, 1 always)
    d.scan_prio = g_xt_scan_prio >= 0 ? g_xt_scan_prio : (2 * H.fast_chunks >= (long long)tile_chunk.back() ? 1 : 0);
    if (G > 1) {
        d.partial = (float*)p->scratch((size_t)G * rows * 4, &rc);
        if (rc) return rc;
        p->carry_bytes = (long long)G * rows * 4;
    }
    d.groups = G;
    d.nblocks = nb * G;
    d.ntiles = (int)tile_slice.size();
    d.nchunks = tile_chunk.back();
    d.cols = cols;
    d.rows = rows;
    p->xt = d;
    return CLSPMV_OK;
}

// Sliced BSR r x r (2 <= r <= 4) with values in `codec` (0: FP32 as stored)
// and, with delta_cols, block columns as int16 steps within a block row when
// every step fits (else int32).  Codec 0 is the FP32 sliced part of
// clspmv_part_bsr method 0 (kernel name bsr<r><r>_sliced[_d16][_xw]).
int build_bsr_coded(clspmv_ctx_t ctx, clspmv_part_t p, const clspmv_bsr_matrix* m, int codec, int delta_cols) {
    {
        const unsigned bw = m->bsr_bwidth, bh = m->bsr_bheight;
        if (bw != bh || bw < 2 || bw > 4) return CLSPMV_E_UNSUPPORTED;
        const int rows = m->matinfo.height, cols = m->matinfo.width;
        if (cols < (int)bw) {
            set_error("mixed BSR needs at least bw columns");
            return CLSPMV_E_UNSUPPORTED;
        }
        const int brows = m->bsr_row_num, nb = m->bsr_block_num, stride = m->bsr_block_stride;
        if (brows != (rows + (int)bh - 1) / (int)bh || stride < nb || m->bsr_row_ptr[0] != 0 ||
            m->bsr_row_ptr[brows] != nb)
            return CLSPMV_E_INVALID;
        for (int r = 0; r < brows; ++r)
            if (m->bsr_row_ptr[r + 1] < m->bsr_row_ptr[r]) return CLSPMV_E_INVALID;
        if (!cols_ok(m->bsr_col_id, (size_t)nb, 0, std::max(1, cols))) {
            set_error("BSR block column out of range");
            return CLSPMV_E_INVALID;
        }
        const unsigned bs = bw * bh;
        const int S = kBsrSlice, ns = (brows + S - 1) / S;
        std::vector<int> sp((size_t)ns + 1, 0);
        for (int s = 0; s < ns; ++s) {
            int wmax = 0;
            for (int r = s * S; r < std::min(brows, s * S + S); ++r)
                wmax = std::max(wmax, m->bsr_row_ptr[r + 1] - m->bsr_row_ptr[r]);
            if ((int64_t)sp[s] + (int64_t)wmax * S > INT32_MAX) return CLSPMV_E_TOO_LARGE;
            sp[s + 1] = sp[s] + wmax * S;
        }
        const size_t slots = (size_t)sp[ns], sstride = std::max<size_t>(64, slots);
        std::vector<int> scol(slots, 0), base((size_t)std::max(1, brows), 0);
        std::vector<float> sdat(sstride * bs, 0.0f);
        const std::vector<int> fill = bsr_slice_fill(m, brows);
        for (int r = 0; r < brows; ++r) {
            const int s = r / S, t = r % S, width = (sp[s + 1] - sp[s]) / S;
            const int a = m->bsr_row_ptr[r], e = m->bsr_row_ptr[r + 1];
            int last = a < e ? m->bsr_col_id[a] : fill[(size_t)s];
            base[(size_t)r] = last;
            for (int k = 0; k < width; ++k) {
                const size_t idx = (size_t)sp[s] + (size_t)k * S + t;
                if (a + k < e) {
                    last = m->bsr_col_id[a + k];
                    for (unsigned q = 0; q < bs; ++q) sdat[(size_t)q * sstride + idx] = m->bsr_data[(size_t)q * stride + a + k];
                }
                scol[idx] = last;
            }
        }
        // int16 deltas when every step fits, else int32 columns
        bool delta = delta_cols != 0;
        std::vector<int16_t> dcol;
        if (delta) {
            dcol.assign(slots, 0);
            for (int r = 0; r < brows && delta; ++r) {
                const int s = r / S, t = r % S, width = (sp[s + 1] - sp[s]) / S;
                int prev = base[(size_t)r];
                for (int k = 0; k < width; ++k) {
                    const size_t idx = (size_t)sp[s] + (size_t)k * S + t;
                    const int dlt = scol[idx] - prev;
                    if (dlt < INT16_MIN || dlt > INT16_MAX) {
                        delta = false;
                        break;
                    }
                    dcol[idx] = (int16_t)dlt;
                    prev = scol[idx];
                }
            }
        }
        const int nsl = ns;
        std::vector<float> scale;
        std::vector<unsigned char> enc;
        const size_t nv = sstride * bs;
        double bound = 0.0;
        if (codec == 0) {
            enc.assign(reinterpret_cast<unsigned char*>(sdat.data()), reinterpret_cast<unsigned char*>(sdat.data()) + nv * 4);
        } else if (codec == kCodecBf16x2) {
            std::vector<uint32_t> w(nv);
            for (size_t i = 0; i < nv; ++i) {
                w[i] = bf16x2_word(sdat[i]);
                bound = std::max(bound, rel_err(sdat[i], bf16x2_decode(w[i])));
            }
            enc.assign(reinterpret_cast<unsigned char*>(w.data()), reinterpret_cast<unsigned char*>(w.data()) + nv * 4);
        } else if (codec_row_scaled(codec)) {
            scale.assign((size_t)std::max(1, brows), 1.0f);
            for (int r = 0; r < brows; ++r) {
                const int s = r / S, t = r % S, width = (sp[s + 1] - sp[s]) / S;
                float amax = 0.f;
                for (int k = 0; k < width; ++k)
                    for (unsigned q = 0; q < bs; ++q)
                        amax = std::max(amax, std::fabs(sdat[(size_t)q * sstride + (size_t)sp[s] + (size_t)k * S + t]));
                scale[(size_t)r] = row_scale(codec, amax);
            }
            if (int e = check_row_scales(codec, scale)) return e;
            std::vector<uint16_t> h(nv, 0);
            for (int r = 0; r < brows; ++r) {
                const int s = r / S, t = r % S, width = (sp[s + 1] - sp[s]) / S;
                for (int k = 0; k < width; ++k)
                    for (unsigned q = 0; q < bs; ++q) {
                        const size_t i = (size_t)q * sstride + (size_t)sp[s] + (size_t)k * S + t;
                        h[i] = row_encode(codec, sdat[i] / scale[(size_t)r]);
                        bound = std::max(bound, rel_err(sdat[i], (double)scale[(size_t)r] * row_decode(codec, h[i])));
                    }
            }
            enc.assign(reinterpret_cast<unsigned char*>(h.data()), reinterpret_cast<unsigned char*>(h.data()) + nv * 2);
        } else {
            scale.assign((size_t)bs * nsl, 1.0f);
            enc.assign(nv, 0);
            for (unsigned q = 0; q < bs; ++q)
                for (int s = 0; s < ns; ++s) {
                    float amax = 0.f;
                    for (int i = sp[s]; i < sp[s + 1]; ++i) amax = std::max(amax, std::fabs(sdat[(size_t)q * sstride + i]));
                    const float sc = fp8_scale(amax);
                    scale[(size_t)q * nsl + s] = sc;
                    for (int i = sp[s]; i < sp[s + 1]; ++i) {
                        const float a = sdat[(size_t)q * sstride + i];
                        const uint8_t e = f32_to_e4m3(a / sc);
                        enc[(size_t)q * sstride + i] = e;
                        bound = std::max(bound, rel_err(a, (double)sc * e4m3_to_f32(e)));
                    }
                }
        }
        // device order of the values: per 64-slot step of a slice, the bs
        // elements' 64 values one after another (k_mixed.hip: bsr_vpos)
        {
            const size_t esz = enc.size() / nv;
            std::vector<unsigned char> step(enc.size(), 0);
#pragma omp parallel for schedule(static) num_threads(omp_threads())
            for (long long i = 0; i < (long long)slots; ++i)
                for (unsigned q = 0; q < bs; ++q)
                    std::memcpy(&step[(((size_t)i >> 6) * bs + q) * 64 * esz + ((size_t)i & 63) * esz],
                                &enc[((size_t)q * sstride + (size_t)i) * esz], esz);
            enc.swap(step);
        }
        int rc = 0;
        DevBsrMixed d{};
        d.sp = (const int*)p->add(ctx, sp.data(), sp.size() * 4, &rc);
        if (rc) return rc;
        if (delta) {
            d.col = p->add(ctx, dcol.data(), std::max<size_t>(16, slots * 2), &rc);
            if (rc) return rc;
            d.base = (const int*)p->add(ctx, base.data(), base.size() * 4, &rc);
        } else {
            d.col = p->add(ctx, scol.data(), std::max<size_t>(16, slots * 4), &rc);
        }
        if (rc) return rc;
        d.data = p->add(ctx, enc.data(), std::max<size_t>(16, enc.size()), &rc);
        if (rc) return rc;
        if (!scale.empty()) {
            d.scale = (const float*)p->add(ctx, scale.data(), scale.size() * 4, &rc);
            if (rc) return rc;
        }
        d.stride = (int)sstride;
        d.brows = brows;
        d.nsl = nsl;
        d.bh = (int)bh;
        d.bw = (int)bw;
        d.rows = rows;
        d.cols = cols;
        d.codec = codec;
        d.delta = delta ? 1 : 0;
        const std::vector<int2> win = bsr_windows(sp, scol, brows, (int)bw, cols);
        if (!win.empty()) {
            d.win = (const int2*)p->add(ctx, win.data(), win.size() * sizeof(int2), &rc);
            if (rc) return rc;
        }
        p->bsrm = d;
        p->mixed = true;
        char name[48];
        if (codec == 0) {
            std::snprintf(name, sizeof(name), "bsr%u%u_sliced%s%s", bh, bw, delta ? "_d16" : "", d.win ? "_xw" : "");
            fill_info(p, CLSPMV_FMT_BSR, 0, rows, cols, m->matinfo.nnz, name, 1);
        } else {
            std::snprintf(name, sizeof(name), "bsr%u%u_mixed_%s%s%s", bh, bw, codec_name(codec), delta ? "_d16" : "",
                          d.win ? "_xw" : "");
            fill_info(p, CLSPMV_FMT_BSR, kMethodMixed + codec + (delta ? 8 : 0), rows, cols, m->matinfo.nnz, name, 1);
            p->info.codec_bound = bound;
        }
        return CLSPMV_OK;
    }
}

}  // namespace

extern "C" {

int clspmv_part_csr_scalar(clspmv_ctx_t ctx, const clspmv_csr_matrix* m, int method, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        int rc = upload_csr(ctx, p, m);
        if (rc) return rc;
        fill_info(p, CLSPMV_FMT_CSR_SCALAR, 0, m->matinfo.height, m->matinfo.width, m->matinfo.nnz, "csr_scalar", 1);
        (void)method;
        return CLSPMV_OK;
    });
}

// method 0: L lanes per row (L from the mean row length);
// method 1: adaptive row blocks (stream / vector / long-row split).
int clspmv_part_csr_vector(clspmv_ctx_t ctx, const clspmv_csr_matrix* m, int method, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        int rc = upload_csr(ctx, p, m);
        if (rc) return rc;
        const int rows = m->matinfo.height;
        const int* rp = m->csr_row_ptr;
        const long long nnz = rp[rows];
        if (method == kMethodXtile) {
            rc = build_xtile(ctx, p, m->csr_row_ptr, m->csr_col_id, m->csr_data, rows, m->matinfo.width);
            if (rc) return rc;
            for (int i = 0; i < 3; ++i) p->bufs[i].release();  // plain CSR upload unused
            return finish_xtile(p, CLSPMV_FMT_CSR_VECTOR, rows, m->matinfo.width, m->matinfo.nnz);
        }
        if (method == 2) {
            rc = build_csr_sliced(ctx, p, m->csr_row_ptr, m->csr_col_id, m->csr_data, rows, m->matinfo.width);
            if (rc) return rc;
            // the plain CSR upload is not used by this layout
            for (int i = 0; i < 3; ++i) p->bufs[i].release();
            char name[32];
            std::snprintf(name, sizeof(name), "csr_sliced_S%d", p->csrs.nslices);
            fill_info(p, CLSPMV_FMT_CSR_VECTOR, 2, rows, m->matinfo.width, m->matinfo.nnz, name,
                      p->csrs.nslices + (p->csrs.n_long ? 1 : 0));
            p->info.stored_bytes -= p->carry_bytes;
            p->info.alg_bytes -= p->carry_bytes;
            return CLSPMV_OK;
        }
        if (method == 0) {
            double mean = rows ? (double)nnz / rows : 0.0;
            p->csr_lanes = std::min(64, std::max(2, pow2_floor((int)(mean / 2.0 + 0.5))));
            char name[32];
            std::snprintf(name, sizeof(name), "csr_vector_L%d", p->csr_lanes);
            fill_info(p, CLSPMV_FMT_CSR_VECTOR, 0, rows, m->matinfo.width, m->matinfo.nnz, name, 1);
            return CLSPMV_OK;
        }
        std::vector<int4> blocks;
        std::vector<int> long_rows, long_first{0};
        int slots = 0;
        int r = 0;
        while (r < rows) {
            const int len = rp[r + 1] - rp[r];
            if (len > kCsrStreamMaxRowLen) {
                if (len <= kCsrLongChunk) {
                    blocks.push_back(make_int4(r, r + 1, rp[r], rp[r + 1]));
                } else {
                    long_rows.push_back(r);
                    for (int n0 = rp[r]; n0 < rp[r + 1]; n0 += kCsrLongChunk) {
                        int n1 = std::min(rp[r + 1], n0 + kCsrLongChunk);
                        blocks.push_back(make_int4(r, -(slots + 1), n0, n1));
                        ++slots;
                    }
                    long_first.push_back(slots);
                }
                ++r;
                continue;
            }
            const int r0 = r;
            int n = 0;
            while (r < rows && r - r0 < 64) {
                const int l = rp[r + 1] - rp[r];
                if (l > kCsrStreamMaxRowLen || n + l > kCsrStreamMaxNnz) break;
                n += l;
                ++r;
            }
            blocks.push_back(make_int4(r0, r, rp[r0], rp[r]));
        }
        p->csra.csr = p->csr;
        p->csra.nblocks = (int)blocks.size();
        p->csra.blocks = (const int4*)p->add(ctx, blocks.data(), blocks.size() * sizeof(int4), &rc);
        if (rc) return rc;
        p->csra.n_long = (int)long_rows.size();
        p->csra.carry = (double*)p->scratch((size_t)std::max(1, slots) * sizeof(double), &rc);
        if (rc) return rc;
        p->csra.long_rows = (const int*)p->add(ctx, long_rows.data(), long_rows.size() * 4, &rc);
        if (rc) return rc;
        p->csra.long_first = (const int*)p->add(ctx, long_first.data(), long_first.size() * 4, &rc);
        if (rc) return rc;
        fill_info(p, CLSPMV_FMT_CSR_VECTOR, 1, rows, m->matinfo.width, m->matinfo.nnz, "csr_adaptive",
                  p->csra.n_long ? 2 : 1);
        // the carry buffer is scratch, not format storage
        p->info.stored_bytes -= (long long)std::max(1, slots) * 8;
        p->info.alg_bytes -= (long long)std::max(1, slots) * 8;
        return CLSPMV_OK;
    });
}

int clspmv_part_coo(clspmv_ctx_t ctx, const clspmv_coo_matrix* m, int method, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        const int rows = m->matinfo.height, cols = m->matinfo.width, nnz = m->matinfo.nnz;
        if (nnz < 0) return CLSPMV_E_INVALID;
        for (int k = 0; k < nnz; ++k) {
            if (m->coo_row_id[k] < 0 || m->coo_row_id[k] >= rows || m->coo_col_id[k] < 0 || m->coo_col_id[k] >= cols) {
                set_error("COO index out of range at %d", k);
                return CLSPMV_E_INVALID;
            }
            if (k && m->coo_row_id[k] < m->coo_row_id[k - 1]) {
                set_error("COO entries must be sorted by row");
                return CLSPMV_E_INVALID;
            }
        }
        if (method == kMethodXtile || method == 2) {
            std::vector<int> rp((size_t)rows + 1, 0);
            for (int k = 0; k < nnz; ++k) rp[(size_t)m->coo_row_id[k] + 1]++;
            for (int r = 0; r < rows; ++r) rp[(size_t)r + 1] += rp[(size_t)r];
            if (method == 2) {
                int rc = build_csr_sliced(ctx, p, rp.data(), m->coo_col_id, m->coo_data, rows, cols);
                if (rc) return rc;
                return finish_sliced(p, CLSPMV_FMT_COO, rows, cols, nnz, "coo");
            }
            int rc = build_xtile(ctx, p, rp.data(), m->coo_col_id, m->coo_data, rows, cols);
            if (rc) return rc;
            return finish_xtile(p, CLSPMV_FMT_COO, rows, cols, nnz);
        }
        const int chunks = (nnz + kCooChunk - 1) / kCooChunk;
        const size_t padded = (size_t)chunks * kCooChunk;
        std::vector<int> rr(padded), cc(padded);
        std::vector<float> vv(padded, 0.0f);
        if (nnz) {
            std::memcpy(rr.data(), m->coo_row_id, (size_t)nnz * 4);
            std::memcpy(cc.data(), m->coo_col_id, (size_t)nnz * 4);
            std::memcpy(vv.data(), m->coo_data, (size_t)nnz * 4);
            for (size_t k = (size_t)nnz; k < padded; ++k) {
                rr[k] = m->coo_row_id[nnz - 1];
                cc[k] = m->coo_col_id[nnz - 1];
            }
        }
        int rc = 0;
        p->coo.row = (const int*)p->add(ctx, rr.data(), padded * 4, &rc);
        if (rc) return rc;
        p->coo.col = (const int*)p->add(ctx, cc.data(), padded * 4, &rc);
        if (rc) return rc;
        p->coo.val = (const float*)p->add(ctx, vv.data(), padded * 4, &rc);
        if (rc) return rc;
        p->coo.nnz_padded = (int)padded;
        p->coo.nnz = (int)padded;
        p->coo.chunks = chunks;
        p->coo.rows = rows;
        const size_t ncarry = (size_t)std::max(1, 2 * chunks);
        p->coo.carry_row = (int*)p->scratch(ncarry * 4, &rc);
        if (rc) return rc;
        p->coo.carry_val = (double*)p->scratch(ncarry * 8, &rc);
        if (rc) return rc;
        fill_info(p, CLSPMV_FMT_COO, 0, rows, cols, nnz, "coo_segment", 2);
        p->info.stored_bytes -= (long long)ncarry * 12;
        p->info.alg_bytes -= (long long)ncarry * 12;
        (void)method;
        return CLSPMV_OK;
    });
}

int clspmv_part_dia(clspmv_ctx_t ctx, const clspmv_dia_matrix* m, int method, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        const int rows = m->matinfo.height, cols = m->matinfo.width, L = m->dia_length_aligned, D = m->dia_num;
        if (D < 0 || L < rows || rows < 0) return CLSPMV_E_INVALID;
        int rc = 0;
        p->dia.offsets = (const int*)p->add(ctx, m->dia_offsets, (size_t)D * 4, &rc);
        if (rc) return rc;
        p->dia.data = (const float*)p->add(ctx, m->dia_data, (size_t)D * L * 4, &rc);
        if (rc) return rc;
        p->dia.L = L;
        p->dia.ndiag = D;
        p->dia.rows = rows;
        p->dia.cols = cols;
        p->dia.offmin = D ? *std::min_element(m->dia_offsets, m->dia_offsets + D) : 0;
        p->dia.offmax = D ? *std::max_element(m->dia_offsets, m->dia_offsets + D) : 0;
        // reference methods 0 / 2 -> one row per lane, 1 / 3 -> four rows
        // per lane; 4: 2-8 lanes per row (one row per lane at <= 8 diagonals)
        int meth = (method == 1 || method == 3) ? 1 : (method == 4 && D > 8) ? 4 : 0;
        if (meth == 1 && L % 4) meth = 0;
        char name[32];
        if (meth == 4) std::snprintf(name, sizeof(name), "dia_lanes%d", D > 32 ? 8 : D > 16 ? 4 : 2);
        else std::snprintf(name, sizeof(name), "%s", meth ? "dia_row4" : "dia_row");
        fill_info(p, CLSPMV_FMT_DIA, meth, rows, cols, m->matinfo.nnz, name, 1);
        return CLSPMV_OK;
    });
}

int clspmv_part_bdia(clspmv_ctx_t ctx, const clspmv_bdia_matrix* m, int method, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        const int rows = m->matinfo.height, cols = m->matinfo.width, L = m->bdia_length_aligned;
        const int nb = m->bdia_band_num;
        if (nb < 0 || L < rows || m->bdia_bptr[0] != 0) return CLSPMV_E_INVALID;
        int maxw = 0;
        for (int b = 0; b < nb; ++b) {
            int w = m->bdia_bptr[b + 1] - m->bdia_bptr[b];
            if (w < 0) return CLSPMV_E_INVALID;
            maxw = std::max(maxw, w);
        }
        const int D = m->bdia_bptr[nb];
        int rc = 0;
        p->bdia.bptr = (const int*)p->add(ctx, m->bdia_bptr, ((size_t)nb + 1) * 4, &rc);
        if (rc) return rc;
        p->bdia.offsets = (const int*)p->add(ctx, m->bdia_offsets, (size_t)nb * 4, &rc);
        if (rc) return rc;
        p->bdia.data = (const float*)p->add(ctx, m->bdia_data, (size_t)D * L * 4, &rc);
        if (rc) return rc;
        p->bdia.L = L;
        p->bdia.nbands = nb;
        p->bdia.rows = rows;
        p->bdia.cols = cols;
        // methods 1 (gpu_bdia) and 4 (gpu_bdia_v4) stage x in LDS
        int meth = (method == 1 || method == 2 || method == 4) ? 1 : 0;
        if (maxw > 257) meth = 0;
        fill_info(p, CLSPMV_FMT_BDIA, meth, rows, cols, m->matinfo.nnz, meth ? "bdia_lds" : "bdia_row", 1);
        return CLSPMV_OK;
    });
}

static bool block_shape_ok(unsigned bw, unsigned bh) {
    return (bw == 4 || bw == 8) && (bh == 1 || bh == 2 || bh == 4 || bh == 8);
}

int clspmv_part_bell(clspmv_ctx_t ctx, const clspmv_b4ell_matrix* m, int method, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        const unsigned bw = m->b4ell_bwidth, bh = m->b4ell_bheight;
        if (!block_shape_ok(bw, bh)) return CLSPMV_E_UNSUPPORTED;
        const int rows = m->matinfo.height, cols = m->matinfo.width;
        const int brows = m->b4ell_row_num, Hc = m->b4ell_height_aligned, F = m->b4ell_float4_aligned;
        const int slots = m->b4ell_block_num;
        if (Hc < brows || F < 4 * brows || (F % 4) || brows != (rows + (int)bh - 1) / (int)bh) return CLSPMV_E_INVALID;
        const int c4max = (cols + 3) / 4;
        for (int j = 0; j < slots; ++j)
            if (!cols_ok(m->b4ell_col_id + (size_t)j * Hc, (size_t)brows, 0, std::max(1, c4max))) {
                set_error("BELL block column out of range");
                return CLSPMV_E_INVALID;
            }
        int rc = 0;
        p->bell.col = (const int*)p->add(ctx, m->b4ell_col_id, (size_t)Hc * slots * 4, &rc);
        if (rc) return rc;
        p->bell.data = (const float*)p->add(ctx, m->b4ell_data, (size_t)F * bh * (bw / 4) * slots * 4, &rc);
        if (rc) return rc;
        p->bell = DevBell{p->bell.col, p->bell.data, Hc, F, slots, brows, (int)bh, (int)bw / 4, rows, cols};
        char name[32];
        std::snprintf(name, sizeof(name), "bell%u%u", bh, bw);
        fill_info(p, CLSPMV_FMT_BELL, method & 1, rows, cols, m->matinfo.nnz, name, 1);
        return CLSPMV_OK;
    });
}

int clspmv_part_sbell(clspmv_ctx_t ctx, const clspmv_sbell_matrix* m, int method, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        const unsigned bw = m->sbell_bwidth, bh = m->sbell_bheight;
        if (!block_shape_ok(bw, bh)) return CLSPMV_E_UNSUPPORTED;
        const int rows = m->matinfo.height, cols = m->matinfo.width;
        const int S = (int)m->sbell_slice_height, ns = m->sbell_slice_num, brows = m->sbell_row_num;
        if (S <= 0 || brows != (rows + (int)bh - 1) / (int)bh || ns != (brows + S - 1) / S) return CLSPMV_E_INVALID;
        if (m->sbell_slice_ptr[0] != 0) return CLSPMV_E_INVALID;
        const int c4max = std::max(1, (cols + 3) / 4);
        for (int s = 0; s < ns; ++s) {
            int a = m->sbell_slice_ptr[s], b = m->sbell_slice_ptr[s + 1];
            if (b < a || (b - a) % S) return CLSPMV_E_INVALID;
            int w = (b - a) / S;
            for (int r = s * S; r < std::min(brows, s * S + S); ++r)
                for (int j = 0; j < w; ++j) {
                    int c = m->sbell_col_id[(size_t)a + (size_t)j * S + (r - s * S)];
                    if (c < 0 || c >= c4max) {
                        set_error("SBELL block column out of range");
                        return CLSPMV_E_INVALID;
                    }
                }
        }
        const size_t total = (size_t)m->sbell_slice_ptr[ns];
        int rc = 0;
        p->sbell.slice_ptr = (const int*)p->add(ctx, m->sbell_slice_ptr, ((size_t)ns + 1) * 4, &rc);
        if (rc) return rc;
        p->sbell.col = (const int*)p->add(ctx, m->sbell_col_id, total * 4, &rc);
        if (rc) return rc;
        p->sbell.data = (const float*)p->add(ctx, m->sbell_data, total * bw * bh * 4, &rc);
        if (rc) return rc;
        p->sbell.S = S;
        p->sbell.brows = brows;
        p->sbell.bh = (int)bh;
        p->sbell.w4n = (int)bw / 4;
        p->sbell.rows = rows;
        p->sbell.cols = cols;
        char name[32];
        std::snprintf(name, sizeof(name), "sbell%u%u", bh, bw);
        fill_info(p, CLSPMV_FMT_SBELL, 0, rows, cols, m->matinfo.nnz, name, 1);
        (void)method;
        return CLSPMV_OK;
    });
}

int clspmv_part_bcsr(clspmv_ctx_t ctx, const clspmv_b4csr_matrix* m, int method, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        const unsigned bw = m->b4csr_bwidth, bh = m->b4csr_bheight;
        if (!block_shape_ok(bw, bh)) return CLSPMV_E_UNSUPPORTED;
        const int rows = m->matinfo.height, cols = m->matinfo.width;
        const int brows = m->b4csr_row_num, nb = m->b4csr_block_num, N = m->b4csr_aligned_size;
        if (brows != (rows + (int)bh - 1) / (int)bh || N < 4 * nb || (N % 4) || m->b4csr_row_ptr[0] != 0 ||
            m->b4csr_row_ptr[brows] != nb)
            return CLSPMV_E_INVALID;
        for (int r = 0; r < brows; ++r)
            if (m->b4csr_row_ptr[r + 1] < m->b4csr_row_ptr[r]) return CLSPMV_E_INVALID;
        if (!cols_ok(m->b4csr_col_id, (size_t)nb, 0, std::max(1, (cols + 3) / 4))) {
            set_error("BCSR block column out of range");
            return CLSPMV_E_INVALID;
        }
        int rc = 0;
        p->bcsr.row_ptr = (const int*)p->add(ctx, m->b4csr_row_ptr, ((size_t)brows + 1) * 4, &rc);
        if (rc) return rc;
        p->bcsr.col = (const int*)p->add(ctx, m->b4csr_col_id, (size_t)nb * 4, &rc);
        if (rc) return rc;
        p->bcsr.data = (const float*)p->add(ctx, m->b4csr_data, (size_t)N * bh * (bw / 4) * 4, &rc);
        if (rc) return rc;
        double mean = brows ? (double)nb / brows : 0.0;
        int lanes = std::min(16, std::max(2, pow2_floor((int)(mean + 0.5))));
        p->bcsr = DevBcsr{p->bcsr.row_ptr, p->bcsr.col, p->bcsr.data, N / 4, brows, (int)bh, (int)bw / 4, rows, cols, lanes};
        char name[32];
        std::snprintf(name, sizeof(name), "bcsr%u%u_L%d", bh, bw, lanes);
        fill_info(p, CLSPMV_FMT_BCSR, 0, rows, cols, m->matinfo.nnz, name, 1);
        (void)method;
        return CLSPMV_OK;
    });
}

int clspmv_part_bsr(clspmv_ctx_t ctx, const clspmv_bsr_matrix* m, int method, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        const unsigned bw = m->bsr_bwidth, bh = m->bsr_bheight;
        if (bw != bh || bw < 2 || bw > 6) return CLSPMV_E_UNSUPPORTED;
        const int rows = m->matinfo.height, cols = m->matinfo.width;
        const int brows = m->bsr_row_num, nb = m->bsr_block_num, stride = m->bsr_block_stride;
        if (brows != (rows + (int)bh - 1) / (int)bh || stride < nb || m->bsr_row_ptr[0] != 0 ||
            m->bsr_row_ptr[brows] != nb)
            return CLSPMV_E_INVALID;
        for (int r = 0; r < brows; ++r)
            if (m->bsr_row_ptr[r + 1] < m->bsr_row_ptr[r]) return CLSPMV_E_INVALID;
        if (!cols_ok(m->bsr_col_id, (size_t)nb, 0, std::max(1, cols))) {
            set_error("BSR block column out of range");
            return CLSPMV_E_INVALID;
        }
        int rc = 0;
        const unsigned bs = bw * bh;
        if (method == 0 && cols < (int)bw) method = 1;  // the sliced kernel reads bw columns at a time
        // 2x2..4x4: the sliced layout with block columns as int16 steps
        // within a block row when every step fits (half the column bytes)
        if (method == 0 && bw <= 4) return build_bsr_coded(ctx, p, m, 0, 1);
        if (method == 0) {
            // Sliced device layout (see bsr_sliced_kernel): built here from
            // the block-CSR host arrays.
            const int S = kBsrSlice;
            const int ns = (brows + S - 1) / S;
            std::vector<int> sp((size_t)ns + 1, 0);
            for (int s = 0; s < ns; ++s) {
                int wmax = 0;
                for (int r = s * S; r < std::min(brows, s * S + S); ++r)
                    wmax = std::max(wmax, m->bsr_row_ptr[r + 1] - m->bsr_row_ptr[r]);
                if ((int64_t)sp[s] + (int64_t)wmax * S > INT32_MAX) return CLSPMV_E_TOO_LARGE;
                sp[s + 1] = sp[s] + wmax * S;
            }
            const size_t slots = (size_t)sp[ns];
            const size_t sstride = std::max<size_t>(64, slots);
            std::vector<int> scol(slots, 0);
            std::vector<float> sdat(sstride * bs, 0.0f);
            const std::vector<int> fill = bsr_slice_fill(m, brows);
            for (int r = 0; r < brows; ++r) {
                const int s = r / S, t = r % S;
                const int width = (sp[s + 1] - sp[s]) / S;
                const int a = m->bsr_row_ptr[r], e = m->bsr_row_ptr[r + 1];
                int last = fill[(size_t)s];
                for (int k = 0; k < width; ++k) {
                    const size_t idx = (size_t)sp[s] + (size_t)k * S + t;
                    if (a + k < e) {
                        last = m->bsr_col_id[a + k];
                        for (unsigned q = 0; q < bs; ++q)
                            sdat[(size_t)q * sstride + idx] = m->bsr_data[(size_t)q * stride + a + k];
                    }
                    scol[idx] = last;
                }
            }
            p->bsr.row_ptr = (const int*)p->add(ctx, sp.data(), sp.size() * 4, &rc);
            if (rc) return rc;
            p->bsr.col = (const int*)p->add(ctx, scol.data(), slots * 4, &rc);
            if (rc) return rc;
            p->bsr.data = (const float*)p->add(ctx, sdat.data(), sdat.size() * 4, &rc);
            if (rc) return rc;
            const std::vector<int2> win = bsr_windows(sp, scol, brows, (int)bw, cols);
            const int2* dwin = nullptr;
            if (!win.empty()) {
                dwin = (const int2*)p->add(ctx, win.data(), win.size() * sizeof(int2), &rc);
                if (rc) return rc;
            }
            p->bsr = DevBsr{p->bsr.row_ptr, p->bsr.col, p->bsr.data, (int)sstride, brows, (int)bh, (int)bw,
                            rows, cols, 0, dwin};
            char name[32];
            std::snprintf(name, sizeof(name), "bsr%u%u_sliced%s", bh, bw, dwin ? "_xw" : "");
            fill_info(p, CLSPMV_FMT_BSR, 0, rows, cols, m->matinfo.nnz, name, 1);
            return CLSPMV_OK;
        }
        p->bsr.row_ptr = (const int*)p->add(ctx, m->bsr_row_ptr, ((size_t)brows + 1) * 4, &rc);
        if (rc) return rc;
        p->bsr.col = (const int*)p->add(ctx, m->bsr_col_id, (size_t)nb * 4, &rc);
        if (rc) return rc;
        p->bsr.data = (const float*)p->add(ctx, m->bsr_data, (size_t)stride * bs * 4, &rc);
        if (rc) return rc;
        double mean = brows ? (double)nb / brows : 0.0;
        int lanes = (method >= 2 && method <= 16) ? pow2_floor(method) : std::min(16, std::max(2, pow2_floor((int)(mean + 0.5))));
        p->bsr = DevBsr{p->bsr.row_ptr, p->bsr.col, p->bsr.data, stride, brows, (int)bh, (int)bw, rows, cols, lanes};
        char name[32];
        std::snprintf(name, sizeof(name), "bsr%u%u_L%d", bh, bw, lanes);
        fill_info(p, CLSPMV_FMT_BSR, lanes, rows, cols, m->matinfo.nnz, name, 1);
        return CLSPMV_OK;
    });
}

// ---------------------------------------------------------------- mixed precision

extern "C" {

int clspmv_part_dia_mixed(clspmv_ctx_t ctx, const clspmv_dia_matrix* m, int codec, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        if (!codec_ok(codec)) return CLSPMV_E_INVALID;
        const int rows = m->matinfo.height, cols = m->matinfo.width, L = m->dia_length_aligned, nd = m->dia_num;
        if (rows < 0 || L < rows || nd < 0 || (L % 256) != 0) {
            set_error("mixed DIA needs dia_length_aligned a multiple of 256 (coo2dia alignment 256)");
            return CLSPMV_E_INVALID;
        }
        const size_t n = (size_t)nd * L;
        const int nsl = L / 256;  // FP8 slices of 256 rows
        std::vector<float> scale;
        std::vector<unsigned char> enc;
        double bound = 0.0;
        if (codec == kCodecBf16x2) {
            std::vector<uint32_t> w(n);
#pragma omp parallel for schedule(static) num_threads(omp_threads()) reduction(max : bound)
            for (long long i = 0; i < (long long)n; ++i) {
                w[(size_t)i] = bf16x2_word(m->dia_data[i]);
                bound = std::max(bound, rel_err(m->dia_data[i], bf16x2_decode(w[(size_t)i])));
            }
            enc.assign(reinterpret_cast<unsigned char*>(w.data()), reinterpret_cast<unsigned char*>(w.data()) + n * 4);
        } else if (codec_row_scaled(codec)) {
            scale.assign((size_t)L, 1.0f);
            for (int r = 0; r < L; ++r) {
                float a = 0.f;
                for (int d = 0; d < nd; ++d) a = std::max(a, std::fabs(m->dia_data[(size_t)d * L + r]));
                scale[(size_t)r] = row_scale(codec, a);
            }
            if (int e = check_row_scales(codec, scale)) return e;
            std::vector<uint16_t> h(n);
#pragma omp parallel for schedule(static) num_threads(omp_threads()) reduction(max : bound)
            for (long long i = 0; i < (long long)n; ++i) {
                const float sc = scale[(size_t)(i % L)];
                h[(size_t)i] = row_encode(codec, m->dia_data[i] / sc);
                bound = std::max(bound, rel_err(m->dia_data[i], (double)sc * row_decode(codec, h[(size_t)i])));
            }
            enc.assign(reinterpret_cast<unsigned char*>(h.data()), reinterpret_cast<unsigned char*>(h.data()) + n * 2);
        } else {
            scale.assign((size_t)nd * nsl, 1.0f);
            enc.assign(n, 0);
#pragma omp parallel for schedule(static) num_threads(omp_threads()) reduction(max : bound)
            for (long long q = 0; q < (long long)nd * nsl; ++q) {
                const size_t d = (size_t)q / nsl, s = (size_t)q % nsl;
                const float* v = m->dia_data + d * L + s * 256;
                float a = 0.f;
                for (int i = 0; i < 256; ++i) a = std::max(a, std::fabs(v[i]));
                const float sc = fp8_scale(a);
                scale[(size_t)q] = sc;
                for (int i = 0; i < 256; ++i) {
                    const uint8_t e = f32_to_e4m3(v[i] / sc);
                    enc[d * L + s * 256 + (size_t)i] = e;
                    bound = std::max(bound, rel_err(v[i], (double)sc * e4m3_to_f32(e)));
                }
            }
        }
        int rc = 0;
        DevDiaMixed d{};
        d.offsets = (const int*)p->add(ctx, m->dia_offsets, (size_t)std::max(1, nd) * 4, &rc);
        if (rc) return rc;
        d.data = p->add(ctx, enc.data(), std::max<size_t>(16, enc.size()), &rc);
        if (rc) return rc;
        if (!scale.empty()) {
            d.scale = (const float*)p->add(ctx, scale.data(), scale.size() * 4, &rc);
            if (rc) return rc;
        }
        d.L = L;
        d.ndiag = nd;
        d.nsl = nsl;
        d.rows = rows;
        d.cols = cols;
        d.codec = codec;
        d.offmin = nd ? *std::min_element(m->dia_offsets, m->dia_offsets + nd) : 0;
        d.offmax = nd ? *std::max_element(m->dia_offsets, m->dia_offsets + nd) : 0;
        p->diam = d;
        p->mixed = true;
        char name[40];
        std::snprintf(name, sizeof(name), "dia_mixed_%s", codec_name(codec));
        fill_info(p, CLSPMV_FMT_DIA, kMethodMixed + codec, rows, cols, m->matinfo.nnz, name, 1);
        p->info.codec_bound = bound;
        return CLSPMV_OK;
    });
}

int clspmv_part_bsr_mixed(clspmv_ctx_t ctx, const clspmv_bsr_matrix* m, int codec, int delta_cols,
                          clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        if (!codec_ok(codec)) return CLSPMV_E_INVALID;
        return build_bsr_coded(ctx, p, m, codec, delta_cols);
    });
}

int clspmv_part_csr_mixed(clspmv_ctx_t ctx, const clspmv_csr_matrix* m, int codec, clspmv_part_t* out) {
    return make_part(ctx, out, [&](clspmv_part_t p) {
        if (!codec_ok(codec)) return CLSPMV_E_INVALID;
        const int rows = m->matinfo.height, cols = m->matinfo.width;
        const int rc = build_xtile(ctx, p, m->csr_row_ptr, m->csr_col_id, m->csr_data, rows, cols, codec);
        if (rc) return rc;
        p->mixed = true;
        return finish_xtile(p, CLSPMV_FMT_CSR_VECTOR, rows, cols, m->matinfo.nnz);
    });
}

}  // extern "C"

extern "C" int clspmv_cocktail_option(const char* name, long long value);  // cocktail.hip

int clspmv_set_option(const char* name, long long value) {
    if (!name) return CLSPMV_E_INVALID;
    if (std::strncmp(name, "cocktail_", 9) == 0) {
        const int rc = clspmv_cocktail_option(name, value);
        if (rc) set_error("option %s: unknown or value %lld out of range", name, value);
        return rc;
    }
    if (std::strcmp(name, "xtile_max_groups") == 0) {
        if (value < 1 || value > kXtMaxGroups) return CLSPMV_E_INVALID;
        g_xt_max_groups = (int)value;
        return CLSPMV_OK;
    }
    if (std::strcmp(name, "xtile_scan_prio") == 0) {
        if (value < -1 || value > 1) return CLSPMV_E_INVALID;
        g_xt_scan_prio = (int)value;
        return CLSPMV_OK;
    }
    if (std::strcmp(name, "bsr_lds_window") == 0) {
        if (value < 0 || value > 2) return CLSPMV_E_INVALID;
        g_bsr_window = (int)value;
        return CLSPMV_OK;
    }
    if (std::strcmp(name, "xtile_blocks") == 0) {
        if (value < 1 || value > (1 << 20)) return CLSPMV_E_INVALID;
        g_xt_blocks = (int)value;
        return CLSPMV_OK;
    }
    if (std::strcmp(name, "col_slice_bytes") == 0) {
        if (value < 256) return CLSPMV_E_INVALID;
        g_slice_x_bytes = value;
        return CLSPMV_OK;
    }
    set_error("unknown option %s", name);
    return CLSPMV_E_INVALID;
}

int clspmv_part_launch(clspmv_part_t p, const float* x, float* y, hipStream_t st) {
    if (p->mixed) {
        if (p->xt.codec) return launch_xtile(p->xt, x, y, st);
        if (p->info.format == CLSPMV_FMT_DIA) return launch_dia_mixed(p->diam, x, y, st);
        if (p->info.format == CLSPMV_FMT_BSR) return launch_bsr_mixed(p->bsrm, x, y, st);
        return (int)hipErrorInvalidValue;
    }
    if (p->info.method == kMethodXtile &&
        (p->info.format == CLSPMV_FMT_ELL || p->info.format == CLSPMV_FMT_SELL ||
         p->info.format == CLSPMV_FMT_CSR_VECTOR || p->info.format == CLSPMV_FMT_COO))
        return launch_xtile(p->xt, x, y, st);
    switch (p->info.format) {
        case CLSPMV_FMT_ELL:
            return p->info.method == 2 ? launch_csr_sliced(p->csrs, x, y, st)
                                       : launch_ell(p->ell, p->info.method, x, y, st);
        case CLSPMV_FMT_SELL:
            return p->info.method == 2 ? launch_csr_sliced(p->csrs, x, y, st) : launch_sell(p->sell, x, y, st);
        case CLSPMV_FMT_CSR_SCALAR: return launch_csr_scalar(p->csr, x, y, st);
        case CLSPMV_FMT_CSR_VECTOR:
            return p->info.method == 0   ? launch_csr_vector(p->csr, p->csr_lanes, x, y, st)
                   : p->info.method == 2 ? launch_csr_sliced(p->csrs, x, y, st)
                                         : launch_csr_adaptive(p->csra, x, y, st);
        case CLSPMV_FMT_COO:
            return p->info.method == 2 ? launch_csr_sliced(p->csrs, x, y, st) : launch_coo(p->coo, x, y, st);
        case CLSPMV_FMT_DIA: return launch_dia(p->dia, p->info.method, x, y, st);
        case CLSPMV_FMT_BDIA: return launch_bdia(p->bdia, p->info.method, x, y, st);
        case CLSPMV_FMT_BELL: return launch_bell(p->bell, x, y, st);
        case CLSPMV_FMT_SBELL: return launch_sbell(p->sbell, x, y, st);
        case CLSPMV_FMT_BCSR: return launch_bcsr(p->bcsr, x, y, st);
        case CLSPMV_FMT_BSR: return launch_bsr(p->bsr, x, y, st);
        default: return (int)hipErrorInvalidValue;
    }
}

int clspmv_part_spmv(clspmv_part_t p, const float* d_x, float* d_y, int ntimes) {
    if (!p || !d_x || !d_y || ntimes < 0) return CLSPMV_E_INVALID;
    if ((reinterpret_cast<uintptr_t>(d_x) | reinterpret_cast<uintptr_t>(d_y)) & 15) {
        set_error("x and y must be 16-byte aligned device pointers");
        return CLSPMV_E_INVALID;
    }
    hipStream_t st = p->ctx->stream();
    for (int i = 0; i < ntimes; ++i) {
        int e = clspmv_part_launch(p, d_x, d_y, st);
        if (e) return hip_status((hipError_t)e, "kernel launch");
    }
    return CLSPMV_OK;
}

int clspmv_part_xtile_stats(clspmv_part_t p, long long* tiles, long long* staged_x_bytes) {
    if (!p) return CLSPMV_E_INVALID;
    if (tiles) *tiles = p->xt_tiles;
    if (staged_x_bytes) *staged_x_bytes = p->xt_staged_bytes;
    return CLSPMV_OK;
}

int clspmv_part_get_info(clspmv_part_t p, clspmv_part_info* info) {
    if (!p || !info) return CLSPMV_E_INVALID;
    *info = p->info;
    return CLSPMV_OK;
}

int clspmv_part_free(clspmv_part_t p) {
    if (!p) return CLSPMV_E_INVALID;
    (void)hipStreamSynchronize(p->ctx->stream());
    delete p;
    return CLSPMV_OK;
}

// ---------------------------------------------------------------- allocation-free COO

int clspmv_coo_spmv_noalloc(clspmv_ctx_t ctx, int rows, int nnz, const int* d_row, const int* d_col,
                            const float* d_val, const float* d_x, float* d_y) {
    if (!ctx || rows < 0 || nnz < 0) return CLSPMV_E_INVALID;
    if (nnz == 0) return CLSPMV_OK;
    if (!d_row || !d_col || !d_val || !d_x || !d_y || rows == 0) return CLSPMV_E_INVALID;
    if (nnz > kNoAllocChunks * kCooChunk) {
        set_error("clspmv_coo_spmv_noalloc: nnz %d above the %d its scratch holds", nnz, kNoAllocChunks * kCooChunk);
        return CLSPMV_E_TOO_LARGE;
    }
    if ((reinterpret_cast<uintptr_t>(d_row) | reinterpret_cast<uintptr_t>(d_col) |
         reinterpret_cast<uintptr_t>(d_val)) & 15) {
        set_error("clspmv_coo_spmv_noalloc: row, col and val must be 16-byte aligned device pointers");
        return CLSPMV_E_INVALID;
    }
    void* base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_noalloc_ptr[ctx->device] : nullptr;
    if (!base) {
        set_error("clspmv_coo_spmv_noalloc: scratch of device %d not set up", ctx->device);
        return CLSPMV_E_INVALID;
    }
    DevCoo m{};
    m.row = d_row;
    m.col = d_col;
    m.val = d_val;
    m.chunks = (nnz + kCooChunk - 1) / kCooChunk;
    m.nnz_padded = m.chunks * kCooChunk;
    m.nnz = nnz;
    m.rows = rows;
    // every carry slot of the m.chunks chunks is written by the segment
    // kernel before the carry kernel reads it: nothing to clear
    m.carry_row = static_cast<int*>(base);
    m.carry_val = reinterpret_cast<double*>(static_cast<unsigned char*>(base) + kNoAllocRowBytes);
    return hip_status((hipError_t)launch_coo(m, d_x, d_y, ctx->stream()), "coo_spmv_noalloc launch");
}

// ---------------------------------------------------------------- probes

int clspmv_probe_copy(clspmv_ctx_t ctx, const void* d_src, void* d_dst, size_t bytes, int ntimes, float* gbps) {
    if (!ctx || !gbps || (bytes % 16) || ntimes <= 0) return CLSPMV_E_INVALID;
    hipStream_t st = ctx->stream();
    int e = launch_copy_probe((const float4*)d_src, (float4*)d_dst, bytes / 16, st);
    if (e) return hip_status((hipError_t)e, "copy probe");
    (void)hipEventRecord(ctx->ev0, st);
    for (int i = 0; i < ntimes; ++i) launch_copy_probe((const float4*)d_src, (float4*)d_dst, bytes / 16, st);
    (void)hipEventRecord(ctx->ev1, st);
    hipError_t he = hipEventSynchronize(ctx->ev1);
    if (he != hipSuccess) return hip_status(he, "copy probe sync");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    *gbps = (float)(2.0 * (double)bytes * ntimes / (ms * 1e-3) / 1e9);
    // the copy is checked: every 16-byte word of dst equals src
    DevBuf cnt;
    int rc = cnt.alloc(sizeof(unsigned long long));
    unsigned long long bad = 0;
    if (!rc) rc = clspmv_memcpy_h2d(ctx, cnt.ptr, &bad, sizeof(bad));
    if (!rc) rc = hip_status((hipError_t)launch_compare_probe(d_src, d_dst, bytes / 16, (unsigned long long*)cnt.ptr, st),
                             "compare probe");
    if (!rc) rc = clspmv_memcpy_d2h(ctx, &bad, cnt.ptr, sizeof(bad));
    cnt.release();
    if (!rc && bad) {
        set_error("copy probe: %llu of %zu 16-byte words differ", bad, bytes / 16);
        return CLSPMV_E_HIP;
    }
    return rc;
}

// Launch cost of a captured hipGraph of `nodes` empty kernels: microseconds
// per graph launch over ntimes back-to-back launches.
int clspmv_probe_graph_launch(clspmv_ctx_t ctx, int nodes, int ntimes, float* usec) {
    if (!ctx || !usec || nodes <= 0 || nodes > 4096 || ntimes <= 0) return CLSPMV_E_INVALID;
    hipStream_t cs = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_status(e, "hipStreamCreate");
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        for (int i = 0; i < nodes; ++i) launch_empty(cs);
        e = hipStreamEndCapture(cs, &graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) return hip_status(e, "graph capture");
    hipStream_t st = ctx->stream();
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipGraphLaunch(exec, st);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, st);
    for (int i = 0; i < ntimes && e == hipSuccess; ++i) e = hipGraphLaunch(exec, st);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, st);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    (void)hipGraphExecDestroy(exec);
    if (e != hipSuccess) return hip_status(e, "graph launch probe");
    *usec = ms * 1000.f / ntimes;
    return CLSPMV_OK;
}

int clspmv_probe_launch(clspmv_ctx_t ctx, int ntimes, float* usec) {
    if (!ctx || !usec || ntimes <= 0) return CLSPMV_E_INVALID;
    hipStream_t st = ctx->stream();
    launch_empty(st);
    (void)hipEventRecord(ctx->ev0, st);
    for (int i = 0; i < ntimes; ++i) launch_empty(st);
    (void)hipEventRecord(ctx->ev1, st);
    hipError_t he = hipEventSynchronize(ctx->ev1);
    if (he != hipSuccess) return hip_status(he, "launch probe sync");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    *usec = ms * 1000.f / ntimes;
    return CLSPMV_OK;
}

int clspmv_probe_launch_grid(clspmv_ctx_t ctx, int groups, int threads, int ntimes, float* usec) {
    if (!ctx || !usec || ntimes <= 0 || groups <= 0 || threads <= 0 || threads > 1024) return CLSPMV_E_INVALID;
    hipStream_t st = ctx->stream();
    for (int i = 0; i < 10; ++i) launch_empty(st, groups, threads);
    (void)hipEventRecord(ctx->ev0, st);
    for (int i = 0; i < ntimes; ++i) launch_empty(st, groups, threads);
    (void)hipEventRecord(ctx->ev1, st);
    hipError_t he = hipEventSynchronize(ctx->ev1);
    if (he != hipSuccess) return hip_status(he, "launch probe sync");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    *usec = ms * 1000.f / ntimes;
    return CLSPMV_OK;
}

}  // extern "C"

extern "C" {

// Host arrays of a CSR matrix's LDS x-tiled layout (xt_host_layout), for
// tests and layout inspection; no device work.
int clspmv_xtile_layout_csr(const clspmv_csr_matrix* m, clspmv_xtile_layout* out) {
    if (!m || !out) return CLSPMV_E_INVALID;
    *out = clspmv_xtile_layout{};
    if (int rc = validate_csr(m)) return rc;
    XtHostLayout H;
    if (int rc = xt_host_layout(m->csr_row_ptr, m->csr_col_id, m->csr_data, m->matinfo.height, m->matinfo.width, 0,
                                &H))
        return rc;
    auto copy = [](const auto& v, auto** dst) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        *dst = (T*)std::malloc(std::max<size_t>(1, v.size()) * sizeof(T));
        if (*dst && !v.empty()) std::memcpy(*dst, v.data(), v.size() * sizeof(T));
        return *dst != nullptr;
    };
    out->nblocks = H.nb;
    out->groups = H.G;
    out->ntiles = (int)H.tile_slice.size();
    out->nchunks = H.tile_chunk.back();
    const bool ok = copy(H.brow, &out->blk_row) && copy(H.blk_tile, &out->blk_tile) &&
                    copy(H.tile_slice, &out->tile_slice) && copy(H.tile_chunk, &out->tile_chunk) &&
                    copy(H.idx, &out->idx) && copy(H.val, &out->val);
    if (!ok) {
        clspmv_free_xtile_layout(out);
        return CLSPMV_E_NOMEM;
    }
    return CLSPMV_OK;
}

void clspmv_free_xtile_layout(clspmv_xtile_layout* l) {
    if (!l) return;
    std::free(l->blk_row);
    std::free(l->blk_tile);
    std::free(l->tile_slice);
    std::free(l->tile_chunk);
    std::free(l->idx);
    std::free(l->val);
    *l = clspmv_xtile_layout{};
}

}  // extern "C"
