// Cocktail GPU driver: the device side of a Cocktail partition.
//
// Replaces cocktail_gpu / cocktail_kernels and their functions
// (A/src/spmv_cocktail.cpp:17-1618).  Each active part of the partition is
// uploaded as a device part (runtime.hip) with the kernel variant that the
// reference's method number maps to; x and y are owned by the object, padded
// as the reference pads them (x to a multiple of 8, y to a multiple of 1024,
// A/src/spmv_cocktail.cpp:230-250).  do_spmv enqueues the parts in the
// reference's fixed order on one stream, so parts that accumulate into the
// same y are serialised as on the reference's in-order queue; the row runs
// of a plan, which write rows of their own, run side by side on up to four
// streams (g_streams).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "clspmv.h"
#include "clspmv_cocktail.h"
#include "host/internal.hpp"
#include "runtime.hpp"

using namespace clspmv;

struct clspmv_gpumat_s {
    clspmv_ctx_t ctx = nullptr;
    DevBuf x, y;
    int width = 0, height = 0;
    std::vector<clspmv_part_t> parts;  // in do_spmv order
    std::vector<int> row0;             // first row of each part (row-range parts of the flat rows)
    // captured graphs: key = passes per graph * 2 + (1: each pass followed
    // by x := y, the iteration of A/src/eval.cpp:19-72)
    std::map<int, hipGraphExec_t> graphs;
    // every part covers rows of its own (row runs of the plan): parts may
    // then run side by side on the extra streams below
    bool disjoint = false;
    std::vector<hipStream_t> aux;  // created on first concurrent call
    std::vector<hipEvent_t> ev;    // ev[0]: fork; ev[1 + a]: join of aux[a]
    ~clspmv_gpumat_s() {
        for (auto s : aux) (void)hipStreamSynchronize(s);
        for (auto s : aux) (void)hipStreamDestroy(s);
        for (auto e : ev) (void)hipEventDestroy(e);
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        for (auto p : parts) clspmv_part_free(p);
        x.release();
        y.release();
    }
};

namespace {

constexpr int kXtile = 4;  // method number of the x-tiled layout

// do_spmv and power_iterate launch each part from the host by default; with
// "cocktail_graph" 1 they replay captured hipGraphs of up to kGraphPasses
// passes (one pass = every part's launch in order).  On MI355X with ROCm
// 7.2 a graph replay costs about 17-27 us per launch and 1.65 us per node at
// 64 nodes, against 1.5-1.65 us per plain launch, so replaying saves nothing
// and costs ~20 us when a call runs few passes.
bool g_use_graph = false;
constexpr int kGraphPasses = 64;

// do_spmv runs the parts of a plan made only of row runs (disjoint rows of
// y) side by side: part j on stream j mod kStreams (stream 0 is the
// context's, the others default-priority streams of the object), forked
// from and joined back to the context stream by events once per call.
// Every pass of a part stays on its stream, in order, and the parts only
// read x, so passes need no join.  A small matrix's parts are each too
// short to fill the GPU, and one after another their launch latencies
// add up.  Switched off with clspmv_set_option("cocktail_streams", 0);
// power_iterate and graph replay stay on one stream.
bool g_streams = true;
constexpr int kStreams = 4;

// x[i] = y[i] for i < n: the vector update of one iteration.
__global__ void copy_vec_kernel(const float* __restrict__ src, float* __restrict__ dst, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Enqueues `passes` passes on stream st, each followed by x := y when
// `iterate`.
int enqueue_passes(clspmv_gpumat_s* g, int passes, bool iterate, hipStream_t st) {
    const float* x = (const float*)g->x.ptr;
    float* y = (float*)g->y.ptr;
    const int n = std::min(g->width, g->height);
    for (int i = 0; i < passes; ++i) {
        for (size_t j = 0; j < g->parts.size(); ++j) {
            const int e = clspmv_part_launch(g->parts[j], x, y + g->row0[j], st);
            if (e) return hip_status((hipError_t)e, "do_spmv launch");
        }
        if (iterate && n > 0) {
            hipLaunchKernelGGL(copy_vec_kernel, dim3((n + 255) / 256), dim3(256), 0, st, y, (float*)g->x.ptr, n);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return hip_status(e, "copy launch");
        }
    }
    return CLSPMV_OK;
}

// The executable graph of `passes` passes (captured on first use, on a
// stream of its own, so that the caller's stream is never in capture mode).
int pass_graph(clspmv_gpumat_s* g, int passes, bool iterate, hipGraphExec_t* out) {
    const int key = passes * 2 + (iterate ? 1 : 0);
    auto it = g->graphs.find(key);
    if (it != g->graphs.end()) {
        *out = it->second;
        return CLSPMV_OK;
    }
    hipStream_t cs = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_status(e, "hipStreamCreate");
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int rc = e == hipSuccess ? CLSPMV_OK : hip_status(e, "hipStreamBeginCapture");
    hipGraph_t graph = nullptr;
    if (!rc) {
        rc = enqueue_passes(g, passes, iterate, cs);
        e = hipStreamEndCapture(cs, &graph);
        if (!rc && e != hipSuccess) rc = hip_status(e, "hipStreamEndCapture");
    }
    hipGraphExec_t exec = nullptr;
    if (!rc) {
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = hip_status(e, "hipGraphInstantiate");
    }
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cs);
    if (rc) return rc;
    g->graphs[key] = exec;
    *out = exec;
    return CLSPMV_OK;
}

// `passes` passes with part j on stream j mod kStreams (see g_streams).
int enqueue_concurrent(clspmv_gpumat_s* g, int passes, hipStream_t st) {
    const int ns = (int)std::min<size_t>(kStreams, g->parts.size());
    while ((int)g->aux.size() < ns - 1) {
        hipStream_t s = nullptr;
        hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e != hipSuccess) return hip_status(e, "hipStreamCreate");
        g->aux.push_back(s);
    }
    while ((int)g->ev.size() < ns) {
        hipEvent_t ev = nullptr;
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return hip_status(e, "hipEventCreate");
        g->ev.push_back(ev);
    }
    hipError_t e = hipEventRecord(g->ev[0], st);
    for (int a = 0; a < ns - 1 && e == hipSuccess; ++a) e = hipStreamWaitEvent(g->aux[(size_t)a], g->ev[0], 0);
    if (e != hipSuccess) return hip_status(e, "fork");
    const float* x = (const float*)g->x.ptr;
    float* y = (float*)g->y.ptr;
    for (int i = 0; i < passes; ++i)
        for (size_t j = 0; j < g->parts.size(); ++j) {
            const int k = (int)(j % (size_t)ns);
            const int r = clspmv_part_launch(g->parts[j], x, y + g->row0[j], k ? g->aux[(size_t)k - 1] : st);
            if (r) return hip_status((hipError_t)r, "do_spmv launch");
        }
    for (int a = 0; a < ns - 1 && e == hipSuccess; ++a) {
        e = hipEventRecord(g->ev[(size_t)a + 1], g->aux[(size_t)a]);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, g->ev[(size_t)a + 1], 0);
    }
    return e == hipSuccess ? CLSPMV_OK : hip_status(e, "join");
}

// ntimes passes on the context stream: whole graphs of kGraphPasses passes,
// then one graph of the remainder.
int run_passes(clspmv_gpumat_s* g, int ntimes, bool iterate) {
    hipStream_t st = g->ctx->stream();
    if (!g_use_graph && !iterate && g_streams && g->disjoint && g->parts.size() > 1 && ntimes > 0)
        return enqueue_concurrent(g, ntimes, st);
    if (!g_use_graph || g->parts.empty()) return enqueue_passes(g, ntimes, iterate, st);
    for (int done = 0; done < ntimes;) {
        const int n = std::min(kGraphPasses, ntimes - done);
        hipGraphExec_t exec = nullptr;
        int rc = pass_graph(g, n, iterate, &exec);
        if (rc) return rc;
        const hipError_t e = hipGraphLaunch(exec, st);
        if (e != hipSuccess) return hip_status(e, "hipGraphLaunch");
        done += n;
    }
    return CLSPMV_OK;
}

// Layout choice for the row-list parts (ELL, SELL, CSR, COO) when x is wider
// than one XCD's L2 (4 MiB): the x-tiled layout (method 4) when its cost is
// below the gather cost, else the column-sliced layout (method 2).  Costs
// fitted on MI355X (uniform random columns, 1M and 4M rows and columns,
// 1-32 entries per row; Zipf rows): gather ~ 6.2 ps per row and 2 MiB
// column slice + 4.65 ps per entry (4M x 4M x 1: 213 us); x-tiled ~ the
// larger of its x staging (every row block -- at most 32768 rows, at least
// 32 blocks -- stages each 8064-column slice it touches, at ~6.7 TB/s:
// 4M x 4M, 1-16 per row: 323-360 us) and 3 ps per entry, plus the pass
// adding the column groups' partial sums and ~10 us of start-up.
bool wide_x(int cols) { return 4.0 * cols > 4.0 * 1024 * 1024; }
bool xtile_pays(long long rows, long long cols, long long nnz) {
    if (nnz <= 0 || cols < 4) return false;
    const double ns = std::ceil(cols / 8064.0);
    const double nb = std::max(32.0, std::ceil(rows / 32768.0));
    const double groups = std::max(1.0, std::min(8.0, std::floor(256.0 / nb)));
    const double per_block = nnz / nb;
    const double staged = nb * ns * (1.0 - std::exp(-per_block / ns)) * 32256.0;
    const double t_xt = std::max(staged / 6.7e12, 3.0e-12 * nnz) +
                        (groups > 1 ? (groups + 2) * rows * 4.0 / 5e12 : 0) + 10e-6;
    const double t_gather = 6.2e-12 * rows * std::ceil(4.0 * cols / (2.0 * 1024 * 1024)) + 4.65e-12 * nnz;
    return t_xt < t_gather;
}

// Kernel variant for each reference method number.
//  BDIA 0 gpu_bdia_nlvec, 3 gpu_bdia_v4_nlvec -> direct x reads (0);
//       1 gpu_bdia, 2 gpu_bdia_g4, 4 gpu_bdia_v4 -> LDS-staged band (1).
//  DIA  0/2 one row per thread -> 0; 1/3 four rows per thread -> 1.
//  ELL  0 -> one row per lane, 1 -> four rows per lane.
//  SELL, SBELL, BELL, BCSR: one kernel per layout.
//  CSR  both reference methods -> adaptive row blocks (1): it needs no
//       tuned workgroup count and balances any row-length distribution.
//  COO  segmented reduction.
//  Wide x: see xtile_pays above.
// Part `which` (0..8: bdia, dia, sbell, bell, bcsr, sell, ell, csr, coo) of
// the partition as the format it is uploaded as and the method passed to
// its clspmv_part_* call.
void partition_part(int which, const clspmv_cocktail* m, int* fmt, int* meth) {
    auto flat = [&](const clspmv_matinfo& a, int near_meth) {
        return !wide_x(a.width) ? near_meth : xtile_pays(a.height, a.width, a.nnz) ? kXtile : 2;
    };
    switch (which) {
        case 0: *fmt = CLSPMV_FMT_BDIA, *meth = m->bdia_meth_num; break;
        case 1: *fmt = CLSPMV_FMT_DIA, *meth = m->dia_meth_num; break;
        case 2: *fmt = CLSPMV_FMT_SBELL, *meth = m->sbell_meth_num; break;
        case 3: *fmt = CLSPMV_FMT_BELL, *meth = m->bell_meth_num; break;
        case 4: *fmt = CLSPMV_FMT_BCSR, *meth = m->bcsr_meth_num; break;
        case 5: *fmt = CLSPMV_FMT_SELL, *meth = flat(m->sell.matinfo, m->sell_meth_num); break;
        case 6: *fmt = CLSPMV_FMT_ELL, *meth = flat(m->ell.matinfo, m->ell_meth_num == 1 ? 1 : 0); break;
        case 7: *fmt = CLSPMV_FMT_CSR_VECTOR, *meth = flat(m->csr.matinfo, 1); break;
        default: *fmt = CLSPMV_FMT_COO, *meth = flat(m->coo.matinfo, 0); break;
    }
}

int add_part(clspmv_gpumat_s* g, int which, const clspmv_cocktail* m) {
    clspmv_part_t p = nullptr;
    int fmt = 0, meth = 0, rc = CLSPMV_OK;
    partition_part(which, m, &fmt, &meth);
    switch (which) {
        case 0: rc = clspmv_part_bdia(g->ctx, &m->bdia, meth, &p); break;
        case 1: rc = clspmv_part_dia(g->ctx, &m->dia, meth, &p); break;
        case 2: rc = clspmv_part_sbell(g->ctx, &m->sbell, meth, &p); break;
        case 3: rc = clspmv_part_bell(g->ctx, &m->bell, meth, &p); break;
        case 4: rc = clspmv_part_bcsr(g->ctx, &m->bcsr, meth, &p); break;
        case 5: rc = clspmv_part_sell(g->ctx, &m->sell, meth, &p); break;
        case 6: rc = clspmv_part_ell(g->ctx, &m->ell, meth, &p); break;
        case 7: rc = clspmv_part_csr_vector(g->ctx, &m->csr, meth, &p); break;
        case 8: rc = clspmv_part_coo(g->ctx, &m->coo, meth, &p); break;
        default: return CLSPMV_E_INVALID;
    }
    if (rc) return rc;
    g->parts.push_back(p);
    g->row0.push_back(0);
    return CLSPMV_OK;
}

// The flat parts (ELL, SELL, CSR, COO) of a partition are all row lists of
// the same matrix.  On MI355X each of them is bound by its pass over x (a
// gather or a staged sweep), so they run as one part: their rows are merged
// (each row sorted by column) and laid out by the same rule as a single CSR
// part.  ELL and SELL slots holding exactly 0.0 -- the padding of those
// formats -- are left out (they add 0 * x[col], which changes nothing for
// finite x).  The partition itself is unchanged.  Switched off with
// clspmv_set_option("cocktail_merge_flat", 0).
bool g_merge_flat = true;

// The merged flat rows are then cut into runs of row blocks (32768 rows)
// by where their columns lie: a block is "local" when at least 3/4 of its
// entries are within 65536 columns of the diagonal (their x stays in L2),
// else "scattered" (the x-tiled or column-sliced layout, by the cost rule
// above, on the run's own shape); local blocks are refined by shape (see
// RunClass).  Each run is one part over its rows, launched with y offset to
// its first row.  c2's band rows run as SELL, its 4 x 4 block rows as
// sliced BSR, its Zipf rows x-tiled.  Switched off with
// clspmv_set_option("cocktail_split_local", 0).
bool g_split_local = true;
// Value codec of the row-run parts that have a mixed-precision variant (the
// x-tiled and sliced BSR ones): 0 = FP32 (default) or a CLSPMV_VAL_* code
// ("cocktail_value_codec").  Other parts stay FP32.
int g_run_codec = 0;
// Error guard of the value codec: a run part keeps its codec only when the
// encoding's bound on the componentwise error (clspmv_part_info.codec_bound,
// valid for any x) is at most this; otherwise it is built in FP32.  Set with
// "cocktail_codec_tol" in units of 1e-9 (default 1e-4; 0 = no guard).
double g_codec_tol = 1e-4;

// Keeps mixed part p when its codec bound passes the guard; else frees it
// and reports false (the caller builds the FP32 part instead).
bool codec_passes(clspmv_part_t p) {
    clspmv_part_info inf;
    if (g_codec_tol <= 0 || (clspmv_part_get_info(p, &inf) == CLSPMV_OK && inf.codec_bound <= g_codec_tol))
        return true;
    clspmv_part_free(p);
    return false;
}
constexpr int kRunBlock = 32768;
constexpr long long kLocalWindow = 65536;

int add_flat_part(clspmv_gpumat_s* g, const clspmv_csr_matrix& m, int meth, int row0) {
    clspmv_part_t p = nullptr;
    int rc = clspmv_part_csr_vector(g->ctx, &m, meth, &p);
    if (rc) return rc;
    g->parts.push_back(p);
    g->row0.push_back(row0);
    return CLSPMV_OK;
}

// Returns kOneRun (nothing added) when single_ok is false and the rows form
// a single run of near-equal rows: the caller then keeps the partition's
// own flat part.
constexpr int kOneRun = 1;

// Row-run classes.  Local runs are refined by shape: rows whose entries lie
// on at most 64 diagonals that they fill to >= 2/3 run as DIA (4 bytes per
// slot, x read along the diagonals); rows of dense aligned b x b blocks
// (b = 4 or 2; bsr_fits) as native BSR; rows of near-equal length (SELL
// with 32-row slices pads <= 25 %) as SELL, scattered ones too while x fits
// one XCD's L2; other rows as CSR row blocks.  "Scattered": under 3/4 of
// the entries within min(65536, max(1024, cols / 8)) columns of the
// diagonal.  The layout of each class by the run's size: run_layout.  The classes that
// change the layout of a single-run matrix (DIA, BSR) are the ones
// >= kLocalDia.
enum RunClass : signed char {
    kScattered = 0,
    kLocalCsr = 1,
    kSell = 2,
    kScatteredSell = 3,  // near-equal rows with scattered columns, x within L2
    kLocalDia = 10,
    kLocalBsr2 = 12,
    kLocalBsr3 = 13,
    kLocalBsr4 = 14
};
constexpr int kDiaMaxDiagonals = 64;

// Block statistics of rows [r0, r1) cut into b x b blocks (block rows
// counted from r0, which is a multiple of b): entries in full blocks, all
// entries, blocks, and the slots of the sliced BSR layout (each 64-block-row
// slice padded to its longest block row).
struct BlockStats {
    long long full = 0, n = 0, blocks = 0, slots = 0;
};
BlockStats block_stats(const clspmv_csr_matrix& m, const std::vector<int>& rp, int r0, int r1, int b) {
    BlockStats st;
    std::vector<int> keys;
    long long slice_max = 0;
    int in_slice = 0;
    for (int br = r0; br < r1; br += b) {
        keys.clear();
        for (int r = br; r < std::min(r1, br + b); ++r)
            for (int k = rp[(size_t)r]; k < rp[(size_t)r + 1]; ++k) keys.push_back(m.csr_col_id[k] / b);
        std::sort(keys.begin(), keys.end());
        st.n += (long long)keys.size();
        long long nb = 0;
        for (size_t i = 0; i < keys.size();) {
            size_t j = i;
            while (j < keys.size() && keys[j] == keys[i]) ++j;
            if ((long long)(j - i) == (long long)b * b) st.full += (long long)(j - i);
            ++nb;
            i = j;
        }
        st.blocks += nb;
        slice_max = std::max(slice_max, nb);
        if (++in_slice == 64) {
            st.slots += 64 * slice_max;
            slice_max = 0;
            in_slice = 0;
        }
    }
    if (in_slice) st.slots += in_slice * slice_max;  // a partial slice: its rows' own padding
    return st;
}

// Share of a row block's entries in full b x b blocks.
double full_block_share(const clspmv_csr_matrix& m, const std::vector<int>& rp, int r0, int r1, int b) {
    const BlockStats st = block_stats(m, rp, r0, r1, b);
    return st.n ? (double)st.full / (double)st.n : 0.0;
}

// Rows [r0, r1) as sliced BSR b x b: >= 95 % of the entries in full blocks
// and the slices' padding at most half the blocks (a few long scattered rows
// among block rows would pad whole slices to their length).
bool bsr_fits(const clspmv_csr_matrix& m, const std::vector<int>& rp, int r0, int r1, int b) {
    const BlockStats st = block_stats(m, rp, r0, r1, b);
    return st.n && 100 * st.full >= 95 * st.n && 2 * st.slots <= 3 * st.blocks;
}

signed char classify_block(const clspmv_csr_matrix& m, const std::vector<int>& rp, int r0, int r1) {
    long long n = 0, near = 0;
    const long long win = std::min<long long>(kLocalWindow, std::max(1024, m.matinfo.width / 8));
    for (int r = r0; r < r1; ++r)
        for (int k = rp[(size_t)r]; k < rp[(size_t)r + 1]; ++k) {
            const long long d = (long long)m.csr_col_id[k] - r;
            near += (d <= win && d >= -win);
            ++n;
        }
    if (n == 0) return -1;
    const bool scattered = 4 * near < 3 * n;
    // repeated (row, col) entries (a matrix file may hold them) are summed
    // by the row-list layouts but would collide in a block: no BSR then
    bool repeats = false;
    for (int r = r0; r < r1 && !repeats; ++r)
        for (int k = rp[(size_t)r] + 1; k < rp[(size_t)r + 1]; ++k)
            if (m.csr_col_id[k] == m.csr_col_id[k - 1]) {
                repeats = true;
                break;
            }
    if (!repeats) {
        // diagonals: distinct column - row offsets (at most 64, filled to 2/3)
        std::vector<long long> off;
        off.reserve((size_t)n);
        bool few = true;
        for (int r = r0; r < r1 && few; ++r)
            for (int k = rp[(size_t)r]; k < rp[(size_t)r + 1]; ++k) {
                off.push_back((long long)m.csr_col_id[k] - r);
                if (off.size() % 4096 == 0) {  // keep the distinct ones, stop early past the limit
                    std::sort(off.begin(), off.end());
                    off.erase(std::unique(off.begin(), off.end()), off.end());
                    if ((int)off.size() > kDiaMaxDiagonals) few = false;
                }
            }
        std::sort(off.begin(), off.end());
        off.erase(std::unique(off.begin(), off.end()), off.end());
        if (few && (int)off.size() <= kDiaMaxDiagonals && 3 * n >= 2LL * (long long)off.size() * (r1 - r0))
            return kLocalDia;
        if (bsr_fits(m, rp, r0, r1, 4)) return kLocalBsr4;
        if (bsr_fits(m, rp, r0, r1, 2)) return kLocalBsr2;
    }
    long long slots = 0;
    for (int s = r0; s < r1; s += 32) {
        int w = 0;
        for (int r = s; r < std::min(r1, s + 32); ++r) w = std::max(w, rp[(size_t)r + 1] - rp[(size_t)r]);
        slots += 32LL * w;
    }
    const bool sell = 4 * slots <= 5 * n;
    if (scattered) return sell && !wide_x(m.matinfo.width) ? kScatteredSell : kScattered;
    return sell ? kSell : kLocalCsr;
}

// Layout (format, method) of a row run of class cls with `rows` rows and
// n entries.  Sizes fitted on MI355X against the autotuner's timings of
// 2400 matrices of 2^10..2^20 rows:
//  * scattered rows: x-tiled (method 4) once the run has >= 2^14 entries
//    and >= 5 per row, or >= 2^19 entries, also with x inside L2 (400 Zipf
//    matrices: 2.0 % mean regret between the two layouts, against 7.5 %
//    for the rule fitted to the x-tiled kernel before its chunk fast path);
//    with x wider than L2 by the cost rule (x-tiled or column-sliced); else
//    adaptive rows;
//  * near-equal scattered rows: SELL, x-tiled from 2^21.5 entries and >= 12
//    per row (uniform rows: 1.1-1.8x faster there);
//  * DIA: one row per lane from 1.5M entries, below that 2-8 lanes per row
//    (all of a row's diagonals in one round of loads; small 27-point
//    stencils ran ~1.5x slower one row per lane than 2 x 2 BSR lanes);
//  * BSR: the sliced layout (one lane per block row) from 2^21 entries;
//    below, 2-16 lanes per block row (method 1) with 2 x 2 blocks for 4 x 4
//    ones -- more waves for a small run (4 x 4 block matrices of 2^11..2^16
//    rows: 2 x 2 lanes 1.1-1.4x faster than 4 x 4 lanes, 1.2-1.8x than the
//    sliced 4 x 4 kernel).
void run_layout(signed char cls, int rows, int cols, long long n, int* fmt, int* meth) {
    const double mean = rows ? (double)n / rows : 0.0;
    *fmt = CLSPMV_FMT_CSR_VECTOR;
    *meth = 1;
    switch (cls) {
        case kScattered:
            if (wide_x(cols)) *meth = xtile_pays(rows, cols, n) ? kXtile : 2;
            else if ((n >= (1LL << 14) && mean >= 5.0) || n >= (1LL << 19)) *meth = kXtile;
            break;
        case kScatteredSell:
            if (n >= 3LL << 20 && mean >= 12.0) *meth = kXtile;
            else *fmt = CLSPMV_FMT_SELL, *meth = 0;
            break;
        case kSell: *fmt = CLSPMV_FMT_SELL, *meth = 0; break;
        case kLocalDia: *fmt = CLSPMV_FMT_DIA, *meth = n < (3LL << 19) ? 4 : 0; break;
        case kLocalBsr2:
        case kLocalBsr3:
        case kLocalBsr4:
            *fmt = CLSPMV_FMT_BSR;
            *meth = n >= (1LL << 21) ? 0 : 1;
            break;
        default: break;  // local CSR rows: adaptive row blocks
    }
}

// Block size of a BSR run's part: the class's, 2 for 4 x 4 blocks in lanes.
unsigned run_block(signed char cls, int meth) {
    const unsigned b = cls == kLocalBsr4 ? 4 : cls == kLocalBsr3 ? 3 : 2;
    return meth == 1 && b == 4 ? 2 : b;
}

// One run as a part of its class and method; rows [0, sub.height) of `sub`.
int add_run_part(clspmv_gpumat_s* g, const clspmv_csr_matrix& sub, signed char cls, int fmt, int meth, int row0) {
    const int cols = sub.matinfo.width, rows = sub.matinfo.height;
    const long long n = sub.matinfo.nnz;
    if (fmt == CLSPMV_FMT_CSR_VECTOR) {
        if (meth != kXtile || !g_run_codec) return add_flat_part(g, sub, meth, row0);
        clspmv_part_t p = nullptr;
        const int rc = clspmv_part_csr_mixed(g->ctx, &sub, g_run_codec, &p);
        if (rc) return rc;
        if (!codec_passes(p)) return add_flat_part(g, sub, meth, row0);
        g->parts.push_back(p);
        g->row0.push_back(row0);
        return CLSPMV_OK;
    }
    // SELL / BSR: through the host conversions of a COO copy of the run
    clspmv_coo_matrix coo{};
    int rc = clspmv_alloc_coo(&coo, cols, rows, (int)n);
    if (rc) return rc;
    for (int r = 0; r < rows; ++r)
        for (int k = sub.csr_row_ptr[r]; k < sub.csr_row_ptr[r + 1]; ++k) {
            coo.coo_row_id[k] = r;
            coo.coo_col_id[k] = sub.csr_col_id[k];
            coo.coo_data[k] = sub.csr_data[k];
        }
    clspmv_part_t p = nullptr;
    if (fmt == CLSPMV_FMT_SELL) {
        clspmv_sell_matrix sell{};
        rc = clspmv_coo2sell(&coo, &sell, 32);
        if (!rc) rc = clspmv_part_sell(g->ctx, &sell, 0, &p);
        clspmv_free_sell(&sell);
    } else if (fmt == CLSPMV_FMT_DIA) {
        clspmv_dia_matrix dia{};
        rc = clspmv_coo2dia(&coo, &dia, 32);
        if (!rc) rc = clspmv_part_dia(g->ctx, &dia, meth, &p);
        clspmv_free_dia(&dia);
    } else {
        const unsigned b = run_block(cls, meth);
        clspmv_bsr_matrix bsr{};
        rc = clspmv_coo2bsr(&coo, &bsr, b, b);
        if (!rc && meth == 1) {
            rc = clspmv_part_bsr(g->ctx, &bsr, 1, &p);
        } else if (!rc && g_run_codec) {
            rc = clspmv_part_bsr_mixed(g->ctx, &bsr, g_run_codec, 1, &p);
            if (!rc && !codec_passes(p)) p = nullptr;
        }
        if (!rc && !p) rc = clspmv_part_bsr(g->ctx, &bsr, 0, &p);
        clspmv_free_bsr(&bsr);
    }
    clspmv_free_coo(&coo);
    if (rc) return rc;
    g->parts.push_back(p);
    g->row0.push_back(row0);
    return CLSPMV_OK;
}

}  // namespace

// The plan: which parts a Cocktail partition runs as on this GPU.  Parts
// 0..8 of the partition (bdia, dia, sbell, bell, bcsr, sell, ell, csr, coo)
// keep their format; the flat ones (5..8) are usually replaced by row runs
// of their merged rows (source 9), each with the class and method chosen
// here.  Built on the host (analysis only), uploaded by
// clspmv_init_plan_kernels.
struct PlanPart {
    int source;        // 0..8 partition part, 9 row run
    signed char cls;   // row runs: RunClass
    int format, method;
    int row0, row1;
    long long nnz;
};
struct clspmv_plan_s {
    clspmv_cocktail own{};  // the partition, when the plan owns it (clspmv_analyze_plan)
    bool owns = false;
    const clspmv_cocktail* m = nullptr;
    std::vector<int> rp, ci;  // merged flat rows (CSR), when merged
    std::vector<float> cv;
    std::vector<PlanPart> parts;
    ~clspmv_plan_s() {
        if (owns) clspmv_free_cocktail(&own);
    }
};

namespace {

// Below this many entries a matrix whose rows fall into several runs runs
// as one part of adaptive row blocks (composites of band, block and Zipf
// rows of 2^15..2^22 entries: 3 runs 1.1-2.1x slower than that one part,
// from 2^22 on 0.76-0.97x).  "cocktail_small_plan" sets it (0: never).
long long g_small_plan = 1LL << 22;

// Rows per sub-block when a class boundary is placed: a multiple of the
// 32-row SELL slice and of the block heights 2 and 4.
constexpr int kRefineRows = 128;

// Places each boundary between runs of different classes (found on run
// blocks of rb rows) at kRefineRows granularity: within the run blocks on
// either side, at the split that keeps the most sub-blocks of the left
// run's class before it and of the right run's class after it.  A
// sub-block of neither class goes to the side whose class is not a layout
// class (DIA, BSR): a composite's block rows beside its scattered rows
// would otherwise carry them into a BSR run.  Runs left empty are dropped,
// neighbours of one class joined, and DIA / BSR runs checked whole again
// (reclassified when the moved rows no longer fit).
void refine_cuts(const clspmv_csr_matrix& m, const std::vector<int>& rp, int rb, std::vector<int>& cut,
                 std::vector<signed char>& rc) {
    const int R = kRefineRows;
    for (size_t i = 1; i + 1 < cut.size(); ++i) {
        const signed char A = rc[i - 1], B = rc[i];
        const int lo = std::max(cut[i - 1], cut[i] - rb), hi = std::min(cut[i + 1], cut[i] + rb);
        const int nsub = (hi - lo + R - 1) / R;  // lo and the old cut are multiples of R
        std::vector<signed char> sc((size_t)nsub);
#pragma omp parallel for schedule(dynamic, 4) num_threads(omp_threads())
        for (int j = 0; j < nsub; ++j) sc[(size_t)j] = classify_block(m, rp, lo + j * R, std::min(hi, lo + (j + 1) * R));
        // score(j): sub-blocks of class A in [0, j) + of class B in [j, nsub)
        int right = 0;
        for (int j = 0; j < nsub; ++j) right += sc[(size_t)j] == B;
        const int j0 = (cut[i] - lo) / R;
        int best = -1, bj = j0, left = 0;
        for (int j = 0; j <= nsub; ++j) {
            const int score = left + right;
            bool take = score > best;
            if (score == best) {
                if (B >= kLocalDia && A < kLocalDia) take = true;                            // mixed rows stay out of B
                else if (!(A >= kLocalDia && B < kLocalDia)) take = std::abs(j - j0) < std::abs(bj - j0);  // nearest
            }
            if (take) {
                best = score;
                bj = j;
            }
            if (j < nsub) {
                left += sc[(size_t)j] == A;
                right -= sc[(size_t)j] == B;
            }
        }
        cut[i] = std::min(hi, lo + bj * R);
    }
    std::vector<int> c2{0};
    std::vector<signed char> r2;
    for (size_t i = 0; i + 1 < cut.size(); ++i) {
        if (cut[i + 1] <= cut[i]) continue;
        signed char k = rc[i];
        if (k >= kLocalDia) {
            const signed char again = classify_block(m, rp, cut[i], cut[i + 1]);
            if (again != k) k = again < 0 ? kLocalCsr : again;
        }
        if (!r2.empty() && r2.back() == k) {
            c2.back() = cut[i + 1];
            continue;
        }
        r2.push_back(k);
        c2.push_back(cut[i + 1]);
    }
    cut.swap(c2);
    rc.swap(r2);
}

// Row runs of the merged flat rows (rp, ci, cv of the plan).  Returns
// kOneRun (nothing planned) when single_ok is false and the rows form a
// single run of a class that keeps the partition's own flat parts.
int plan_runs(clspmv_plan_s* P, bool single_ok) {
    const int rows = P->m->mat_height, cols = P->m->mat_width;
    clspmv_csr_matrix merged{};
    merged.matinfo.width = cols;
    merged.matinfo.height = rows;
    merged.matinfo.nnz = rows ? P->rp[(size_t)rows] : 0;
    merged.csr_row_ptr = P->rp.data();
    merged.csr_col_id = P->ci.data();
    merged.csr_data = P->cv.data();
    const std::vector<int>& rp = P->rp;
    auto push_run = [&](signed char cls, int r0, int r1) {
        const long long n = rp[(size_t)r1] - rp[(size_t)r0];
        PlanPart q{9, cls, 0, 0, r0, r1, n};
        run_layout(cls, r1 - r0, cols, n, &q.format, &q.method);
        P->parts.push_back(q);
    };
    // run blocks: 32768 rows, finer (down to 1024) below 1M rows so that a
    // small matrix still has ~32 blocks to tell its regions apart
    int rb = 1024;
    while (rb < kRunBlock && rb * 32 <= rows) rb *= 2;
    const int nbk = (rows + rb - 1) / rb;
    if (!g_split_local || nbk == 0) {
        if (!single_ok) return kOneRun;
        push_run(kScattered, 0, rows);  // one CSR part, by the cost rule (local x: adaptive rows)
        return CLSPMV_OK;
    }
    // 3 x 3 blocks do not align with power-of-two runs: a matrix made of
    // them (its first run local and >= 95 % of all entries in full blocks
    // counted from row 0) runs whole as sliced BSR 3 x 3
    {
        const int r1 = std::min(rows, rb);
        const signed char c0 = classify_block(merged, rp, 0, r1);
        if (c0 > kScattered && c0 < kLocalDia && full_block_share(merged, rp, 0, r1, 3) >= 0.95 &&
            bsr_fits(merged, rp, 0, rows, 3)) {
            push_run(kLocalBsr3, 0, rows);
            return CLSPMV_OK;
        }
    }
    std::vector<signed char> cls((size_t)nbk);
#pragma omp parallel for schedule(dynamic, 1) num_threads(omp_threads())
    for (int b = 0; b < nbk; ++b)
        cls[(size_t)b] = classify_block(merged, rp, b * rb, std::min(rows, (b + 1) * rb));
    for (int b = 0; b < nbk; ++b)  // empty blocks join the previous run
        if (cls[(size_t)b] < 0) cls[(size_t)b] = b ? cls[(size_t)b - 1] : kLocalCsr;
    std::vector<int> cut{0};
    std::vector<signed char> rc{cls[0]};  // class of run i = rows [cut[i], cut[i + 1])
    for (int b = 1; b < nbk; ++b)
        if (cls[(size_t)b] != cls[(size_t)b - 1]) {
            cut.push_back(b * rb);
            rc.push_back(cls[(size_t)b]);
        }
    cut.push_back(rows);
    if (cut.size() == 2 && cls[0] == kSell && !single_ok) {
        // one run of near-equal rows: the partition's own flat part (its
        // ELL or SELL serves such rows as well as a SELL run)
        return kOneRun;
    }
    refine_cuts(merged, rp, rb, cut, rc);
    if (cut.size() > 2 && merged.matinfo.nnz < g_small_plan) {
        // a small matrix runs as one part of adaptive row blocks: its runs'
        // launches, even side by side, cost more than they save
        push_run(kLocalCsr, 0, rows);
        return CLSPMV_OK;
    }
    const size_t first = P->parts.size();
    for (size_t i = 0; i + 1 < cut.size(); ++i) {
        // neighbouring runs whose classes give one layout (scattered and
        // local CSR rows both as adaptive row blocks, say) run as one part,
        // its layout chosen again for the joined size
        const PlanPart* b = P->parts.size() > first ? &P->parts.back() : nullptr;
        int f = 0, m = 0;
        run_layout(rc[i], cut[i + 1] - cut[i], cols, rp[(size_t)cut[i + 1]] - rp[(size_t)cut[i]], &f, &m);
        if (b && b->format == f && b->method == m && (f != CLSPMV_FMT_BSR || run_block(b->cls, m) == run_block(rc[i], m))) {
            const signed char k = b->cls == kScattered || rc[i] == kScattered ? kScattered : b->cls;
            const int r0 = b->row0;
            P->parts.pop_back();
            push_run(k, r0, cut[i + 1]);
            continue;
        }
        push_run(rc[i], cut[i], cut[i + 1]);
    }
    return CLSPMV_OK;
}

// Merges the flat parts' rows into P->rp / ci / cv.
int merge_flat(clspmv_plan_s* P) {
    const clspmv_cocktail* m = P->m;
    const int rows = m->mat_height;
    std::vector<long long> cnt((size_t)rows + 1, 0);
    const clspmv_ell_matrix& e = m->ell;
    const clspmv_sell_matrix& sl = m->sell;
    const clspmv_csr_matrix& c = m->csr;
    const clspmv_coo_matrix& o = m->coo;
    std::vector<int> coo_rp;
    if (m->ifusecoo) {
        coo_rp.assign((size_t)rows + 1, 0);
        for (int k = 0; k < o.matinfo.nnz; ++k) coo_rp[(size_t)o.coo_row_id[k] + 1]++;
        for (int r = 0; r < rows; ++r) coo_rp[(size_t)r + 1] += coo_rp[(size_t)r];
    }
    auto sell_width = [&](int r) {
        const int S = (int)sl.sell_slice_height, sidx = r / S;
        return (sl.sell_slice_ptr[sidx + 1] - sl.sell_slice_ptr[sidx]) / S;
    };
    auto ell_slot = [&](int r, int k) { return (size_t)k * e.ell_height_aligned + (size_t)r; };
    auto sell_slot = [&](int r, int j) {
        const int S = (int)sl.sell_slice_height;
        return (size_t)sl.sell_slice_ptr[r / S] + (size_t)j * S + (size_t)(r % S);
    };
#pragma omp parallel for schedule(static) num_threads(omp_threads())
    for (int r = 0; r < rows; ++r) {
        long long n = 0;
        if (m->ifuseell)
            for (int k = 0; k < e.ell_num; ++k) n += e.ell_data[ell_slot(r, k)] != 0.0f;
        if (m->ifusesell)
            for (int j = 0, w = sell_width(r); j < w; ++j) n += sl.sell_data[sell_slot(r, j)] != 0.0f;
        if (m->ifusecsr) n += c.csr_row_ptr[r + 1] - c.csr_row_ptr[r];
        if (m->ifusecoo) n += coo_rp[(size_t)r + 1] - coo_rp[(size_t)r];
        cnt[(size_t)r + 1] = n;
    }
    for (int r = 0; r < rows; ++r) cnt[(size_t)r + 1] += cnt[(size_t)r];
    if (cnt[(size_t)rows] > INT32_MAX) return CLSPMV_E_TOO_LARGE;
    std::vector<int>& rp = P->rp;
    rp.assign((size_t)rows + 1, 0);
    P->ci.assign((size_t)cnt[(size_t)rows], 0);
    P->cv.assign((size_t)cnt[(size_t)rows], 0.f);
    for (int r = 0; r <= rows; ++r) rp[(size_t)r] = (int)cnt[(size_t)r];
    int* ci = P->ci.data();
    float* cv = P->cv.data();
#pragma omp parallel for schedule(dynamic, 1024) num_threads(omp_threads())
    for (int r = 0; r < rows; ++r) {
        std::vector<std::pair<int, float>> row;
        row.reserve((size_t)(rp[(size_t)r + 1] - rp[(size_t)r]));
        if (m->ifuseell)
            for (int k = 0; k < e.ell_num; ++k) {
                const size_t i = ell_slot(r, k);
                if (e.ell_data[i] != 0.0f) row.push_back({e.ell_col_id[i], e.ell_data[i]});
            }
        if (m->ifusesell)
            for (int j = 0, w = sell_width(r); j < w; ++j) {
                const size_t i = sell_slot(r, j);
                if (sl.sell_data[i] != 0.0f) row.push_back({sl.sell_col_id[i], sl.sell_data[i]});
            }
        if (m->ifusecsr)
            for (int k = c.csr_row_ptr[r]; k < c.csr_row_ptr[r + 1]; ++k) row.push_back({c.csr_col_id[k], c.csr_data[k]});
        if (m->ifusecoo)
            for (int k = coo_rp[(size_t)r]; k < coo_rp[(size_t)r + 1]; ++k) row.push_back({o.coo_col_id[k], o.coo_data[k]});
        std::stable_sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < row.size(); ++k) {
            ci[(size_t)rp[(size_t)r] + k] = row[k].first;
            cv[(size_t)rp[(size_t)r] + k] = row[k].second;
        }
    }
    return CLSPMV_OK;
}

// The plan of partition m (not owned: it must outlive the plan's upload).
int build_plan(const clspmv_cocktail* m, clspmv_plan_s* P) {
    P->m = m;
    P->parts.clear();
    const unsigned char use[9] = {m->ifusebdia, m->ifusedia, m->ifusesbell, m->ifusebell, m->ifusebcsr,
                                  m->ifusesell, m->ifuseell, m->ifusecsr, m->ifusecoo};
    const int nflat = use[5] + use[6] + use[7] + use[8];
    // a single flat part goes through the merge too when it may split into
    // row runs; if it does not, it keeps its own format
    const bool merge = g_merge_flat && (nflat >= 2 || (nflat == 1 && g_split_local));
    const long long nnz_of[9] = {m->bdia.matinfo.nnz, m->dia.matinfo.nnz, m->sbell.matinfo.nnz,
                                 m->bell.matinfo.nnz, m->bcsr.matinfo.nnz, m->sell.matinfo.nnz,
                                 m->ell.matinfo.nnz, m->csr.matinfo.nnz, m->coo.matinfo.nnz};
    auto push_part = [&](int i) {
        int fmt = 0, meth = 0;
        partition_part(i, m, &fmt, &meth);
        P->parts.push_back(PlanPart{i, -1, fmt, meth, 0, m->mat_height, nnz_of[i]});
    };
    for (int i = 0; i < 9; ++i)
        if (use[i] && !(merge && i >= 5)) push_part(i);
    if (!merge) return CLSPMV_OK;
    int rc = merge_flat(P);
    if (!rc) rc = plan_runs(P, nflat >= 2);
    if (rc == kOneRun) {
        P->rp.clear();
        P->ci.clear();
        P->cv.clear();
        for (int i = 5; i < 9; ++i)
            if (use[i]) push_part(i);
        rc = CLSPMV_OK;
    }
    return rc;
}

int upload_plan(clspmv_gpumat_s* g, const clspmv_plan_s* P) {
    for (const PlanPart& q : P->parts) {
        int rc;
        if (q.source < 9) {
            rc = add_part(g, q.source, P->m);
        } else {
            std::vector<int> sub_rp((size_t)(q.row1 - q.row0) + 1);
            for (int r = q.row0; r <= q.row1; ++r) sub_rp[(size_t)(r - q.row0)] = P->rp[(size_t)r] - P->rp[(size_t)q.row0];
            clspmv_csr_matrix sub{};
            sub.matinfo.width = P->m->mat_width;
            sub.matinfo.height = q.row1 - q.row0;
            sub.matinfo.nnz = sub_rp.back();
            sub.csr_row_ptr = sub_rp.data();
            sub.csr_col_id = const_cast<int*>(P->ci.data()) + P->rp[(size_t)q.row0];
            sub.csr_data = const_cast<float*>(P->cv.data()) + P->rp[(size_t)q.row0];
            rc = add_run_part(g, sub, q.cls, q.format, q.method, q.row0);
        }
        if (rc) return rc;
    }
    return CLSPMV_OK;
}

int check_size(clspmv_gpumat_t g, int n, int limit) {
    if (!g || n < 0 || n > limit) {
        set_error("size %d outside [0, %d]", n, limit);
        return CLSPMV_E_INVALID;
    }
    return CLSPMV_OK;
}

int init_from_plan(clspmv_ctx_t ctx, const clspmv_plan_s* P, const float* vec, const float* res, clspmv_gpumat_t* out) {
    const clspmv_cocktail* m = P->m;
    auto* g = new (std::nothrow) clspmv_gpumat_s();
    if (!g) return CLSPMV_E_NOMEM;
    g->ctx = ctx;
    g->width = m->mat_width;
    g->height = m->mat_height;
    const size_t xn = (size_t)clspmv_aligned_length(g->width, 8);
    const size_t yn = (size_t)clspmv_aligned_length(g->height, 1024);
    int rc = g->x.alloc(xn * 4);
    if (!rc) rc = g->y.alloc(yn * 4);
    if (!rc) rc = hip_status(hipMemsetAsync(g->x.ptr, 0, xn * 4, ctx->stream()), "hipMemset");
    if (!rc) rc = hip_status(hipMemsetAsync(g->y.ptr, 0, yn * 4, ctx->stream()), "hipMemset");
    if (!rc && vec) rc = clspmv_memcpy_h2d(ctx, g->x.ptr, vec, (size_t)g->width * 4);
    if (!rc && res) rc = clspmv_memcpy_h2d(ctx, g->y.ptr, res, (size_t)g->height * 4);
    if (!rc) rc = upload_plan(g, P);
    if (!rc) {
        g->disjoint = !P->parts.empty();
        for (const PlanPart& q : P->parts) g->disjoint = g->disjoint && q.source == 9;
    }
    if (!rc) rc = clspmv_synchronize(ctx);
    if (rc) {
        delete g;
        *out = nullptr;
        return rc;
    }
    *out = g;
    return CLSPMV_OK;
}

}  // namespace

extern "C" {

int clspmv_init_mat_kernels(clspmv_ctx_t ctx, const clspmv_cocktail* m, const float* vec, const float* res,
                            int use_tex, clspmv_gpumat_t* out) {
    (void)use_tex;
    if (!ctx || !m || !out) return CLSPMV_E_INVALID;
    clspmv_plan_s plan;
    const int rc = build_plan(m, &plan);
    if (rc) return rc;
    return init_from_plan(ctx, &plan, vec, res, out);
}

int clspmv_plan_cocktail(const clspmv_cocktail* m, clspmv_plan_t* out) {
    if (!m || !out) return CLSPMV_E_INVALID;
    auto* P = new (std::nothrow) clspmv_plan_s();
    if (!P) return CLSPMV_E_NOMEM;
    const int rc = build_plan(m, P);
    if (rc) {
        delete P;
        return rc;
    }
    *out = P;
    return CLSPMV_OK;
}

int clspmv_analyze_plan(clspmv_bench_t bench, clspmv_coo_matrix* coo, float dia_priority_multiplier,
                        int estimate_flat_partial, int estimate_flat_full, clspmv_plan_t* out) {
    if (!bench || !coo || !out) return CLSPMV_E_INVALID;
    auto* P = new (std::nothrow) clspmv_plan_s();
    if (!P) return CLSPMV_E_NOMEM;
    clspmv_init_cocktail(&P->own);
    P->owns = true;
    int rc = clspmv_analyze_matrix(bench, coo, &P->own, dia_priority_multiplier, estimate_flat_partial,
                                   estimate_flat_full, 0);
    if (!rc) rc = build_plan(&P->own, P);
    if (rc) {
        delete P;
        return rc;
    }
    *out = P;
    return CLSPMV_OK;
}

int clspmv_plan_num_parts(clspmv_plan_t P) { return P ? (int)P->parts.size() : 0; }

int clspmv_plan_get_part(clspmv_plan_t P, int i, clspmv_plan_part* out) {
    if (!P || !out || i < 0 || i >= (int)P->parts.size()) return CLSPMV_E_INVALID;
    const PlanPart& q = P->parts[(size_t)i];
    const int block = q.source == 9 && q.format == CLSPMV_FMT_BSR ? (int)run_block(q.cls, q.method) : 0;
    *out = clspmv_plan_part{q.source, q.format, q.method, q.row0, q.row1, q.nnz, block};
    return CLSPMV_OK;
}

const clspmv_cocktail* clspmv_plan_partition(clspmv_plan_t P) { return P ? P->m : nullptr; }

int clspmv_plan_free(clspmv_plan_t P) {
    if (!P) return CLSPMV_E_INVALID;
    delete P;
    return CLSPMV_OK;
}

int clspmv_init_plan_kernels(clspmv_ctx_t ctx, clspmv_plan_t P, const float* vec, const float* res,
                             clspmv_gpumat_t* out) {
    if (!ctx || !P || !out) return CLSPMV_E_INVALID;
    return init_from_plan(ctx, P, vec, res, out);
}

int clspmv_free_mat_kernels(clspmv_gpumat_t g) {
    if (!g) return CLSPMV_E_INVALID;
    (void)clspmv_synchronize(g->ctx);
    delete g;
    return CLSPMV_OK;
}

int clspmv_do_spmv(clspmv_gpumat_t g, int ntimes) {
    if (!g || ntimes < 0) return CLSPMV_E_INVALID;
    return run_passes(g, ntimes, false);
}

int clspmv_power_iterate(clspmv_gpumat_t g, int iterations) {
    if (!g || iterations < 0) return CLSPMV_E_INVALID;
    if (g->width != g->height) {
        set_error("power_iterate needs a square matrix (%d x %d)", g->height, g->width);
        return CLSPMV_E_INVALID;
    }
    return run_passes(g, iterations, true);
}

int clspmv_cpy_vector_from_cpu(clspmv_gpumat_t g, const float* vec, int n) {
    int rc = check_size(g, n, g ? g->width : 0);
    return rc ? rc : clspmv_memcpy_h2d(g->ctx, g->x.ptr, vec, (size_t)n * 4);
}
int clspmv_cpy_vector_from_gpu(clspmv_gpumat_t g, const void* d_vec, int n) {
    int rc = check_size(g, n, g ? g->width : 0);
    return rc ? rc : clspmv_memcpy_d2d(g->ctx, g->x.ptr, d_vec, (size_t)n * 4);
}
int clspmv_cpy_result_from_cpu(clspmv_gpumat_t g, const float* res, int n) {
    int rc = check_size(g, n, g ? g->height : 0);
    return rc ? rc : clspmv_memcpy_h2d(g->ctx, g->y.ptr, res, (size_t)n * 4);
}
int clspmv_cpy_result_from_gpu(clspmv_gpumat_t g, const void* d_res, int n) {
    int rc = check_size(g, n, g ? g->height : 0);
    return rc ? rc : clspmv_memcpy_d2d(g->ctx, g->y.ptr, d_res, (size_t)n * 4);
}
int clspmv_cpy_result_to_cpu(clspmv_gpumat_t g, float* res, int n) {
    int rc = check_size(g, n, g ? g->height : 0);
    return rc ? rc : clspmv_memcpy_d2h(g->ctx, res, g->y.ptr, (size_t)n * 4);
}
int clspmv_cpy_result_to_gpu(clspmv_gpumat_t g, void* d_res, int n) {
    int rc = check_size(g, n, g ? g->height : 0);
    return rc ? rc : clspmv_memcpy_d2d(g->ctx, d_res, g->y.ptr, (size_t)n * 4);
}

int clspmv_gpumat_buffers(clspmv_gpumat_t g, void** d_vec, void** d_res) {
    if (!g) return CLSPMV_E_INVALID;
    if (d_vec) *d_vec = g->x.ptr;
    if (d_res) *d_res = g->y.ptr;
    return CLSPMV_OK;
}

int clspmv_gpumat_num_parts(clspmv_gpumat_t g) { return g ? (int)g->parts.size() : 0; }

// evaluate (A/src/spmv_cocktail.cpp:1431-1455 with check_and_time
// :1134-1154): upload the partition with x = vec and y = res, one checked
// do_spmv against coores (the caller's CPU y = res + A vec), then ntimes
// timed do_spmv.  The check reports the reference's two_vec_compare fields
// (A/src/util.cpp:93-131: Euclidean distance, and the largest difference of
// magnitudes and its ratio to the larger magnitude, sign-blind) and a signed
// one, max |a - b| / max(|a|, |b|).
int clspmv_evaluate_report(clspmv_ctx_t ctx, const clspmv_cocktail* mat, const float* vec, const float* res,
                           const float* coores, int ntimes, int verbose, clspmv_eval_report* out) {
    if (!ctx || !mat || !vec || !res || !coores || ntimes < 1 || !out) return CLSPMV_E_INVALID;
    *out = clspmv_eval_report{};
    clspmv_gpumat_t g = nullptr;
    int rc = clspmv_init_mat_kernels(ctx, mat, vec, res, 0, &g);
    if (rc) return rc;
    const int h = mat->mat_height;
    std::vector<float> y((size_t)std::max(1, h));
    rc = clspmv_cpy_result_from_cpu(g, res, h);
    if (!rc) rc = clspmv_do_spmv(g, 1);
    if (!rc) rc = clspmv_cpy_result_to_cpu(g, y.data(), h);
    if (!rc) {
        double sum = 0.0;
        for (int i = 0; i < h; ++i) {
            const double a = coores[i], b = y[(size_t)i];
            sum += (a - b) * (a - b);
            const double d = std::fabs(std::fabs(a) - std::fabs(b)), mx = std::max(std::fabs(a), std::fabs(b));
            if (d > out->max_diff) {
                out->max_diff = d;
                out->max_diff_id = i;
            }
            if (mx > 0) {
                out->max_diff_ratio = std::max(out->max_diff_ratio, d / mx);
                out->max_rel_err = std::max(out->max_rel_err, std::fabs(a - b) / mx);
            }
        }
        out->distance = std::sqrt(sum);
    }
    float ms = 0.f;
    if (!rc) rc = clspmv_synchronize(ctx);
    if (!rc) rc = clspmv_event_elapsed_begin(ctx);
    if (!rc) rc = clspmv_do_spmv(g, ntimes);
    if (!rc) rc = clspmv_event_elapsed_end(ctx, &ms);
    if (!rc) {
        long long alg = 4LL * mat->mat_width + 8LL * h;
        for (size_t i = 0; i < g->parts.size(); ++i) {
            clspmv_part_info inf;
            if (clspmv_part_get_info(g->parts[i], &inf) == CLSPMV_OK) alg += inf.stored_bytes;
        }
        out->ntimes = ntimes;
        out->nparts = (int)g->parts.size();
        out->ms_per_spmv = (double)ms / ntimes;
        out->gflops = 2.0 * mat->mat_nnz / (out->ms_per_spmv * 1e-3) / 1e9;
        out->alg_gbps = (double)alg / (out->ms_per_spmv * 1e-3) / 1e9;
        if (verbose) {
            std::printf("Max diff id %d coo res %f res %f \n", out->max_diff_id, h ? coores[out->max_diff_id] : 0.f,
                        h ? y[(size_t)out->max_diff_id] : 0.f);
            std::printf("\nCorrectness Check: Distance %e max diff %e max diff ratio %e vec size %d\n", out->distance,
                        out->max_diff, out->max_diff_ratio, h);
            std::printf("\nCocktail gpu time %lf ms GFLOPS %lf\n\n", out->ms_per_spmv, out->gflops);
        }
    }
    const int rf = clspmv_free_mat_kernels(g);
    return rc ? rc : rf;
}

int clspmv_evaluate(const clspmv_cocktail* mat, const float* vec, const float* res, const float* coores) {
    clspmv_ctx_t ctx = nullptr;
    int rc = clspmv_init(0, &ctx);
    if (rc) return rc;
    clspmv_eval_report rep;
    rc = clspmv_evaluate_report(ctx, mat, vec, res, coores, 1000, 1, &rep);
    const int rf = clspmv_finalize(ctx);
    return rc ? rc : rf;
}

__attribute__((visibility("hidden"))) int clspmv_cocktail_option(const char* name, long long value) {
    if (name && std::strcmp(name, "cocktail_merge_flat") == 0) {
        g_merge_flat = value != 0;
        return CLSPMV_OK;
    }
    if (name && std::strcmp(name, "cocktail_codec_tol") == 0) {
        if (value < 0) return CLSPMV_E_INVALID;
        g_codec_tol = (double)value * 1e-9;
        return CLSPMV_OK;
    }
    if (name && std::strcmp(name, "cocktail_graph") == 0) {
        g_use_graph = value != 0;
        return CLSPMV_OK;
    }
    if (name && std::strcmp(name, "cocktail_small_plan") == 0) {
        if (value < 0) return CLSPMV_E_INVALID;
        g_small_plan = value;
        return CLSPMV_OK;
    }
    if (name && std::strcmp(name, "cocktail_streams") == 0) {
        g_streams = value != 0;
        return CLSPMV_OK;
    }
    if (name && std::strcmp(name, "cocktail_split_local") == 0) {
        g_split_local = value != 0;
        return CLSPMV_OK;
    }
    if (name && std::strcmp(name, "cocktail_value_codec") == 0) {
        if (value < 0 || value > CLSPMV_VAL_I16_ROW) return CLSPMV_E_INVALID;
        g_run_codec = (int)value;
        return CLSPMV_OK;
    }
    return CLSPMV_E_INVALID;
}

int clspmv_gpumat_part_info(clspmv_gpumat_t g, int i, clspmv_part_info* info) {
    if (!g || i < 0 || i >= (int)g->parts.size()) return CLSPMV_E_INVALID;
    return clspmv_part_get_info(g->parts[(size_t)i], info);
}

}  // extern "C"
