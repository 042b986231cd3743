// LDS x-tiled SpMV (y += A*x) for matrices whose columns are scattered over a
// wide x (uniform random columns, power-law rows).
//
// A random-column SpMV that reads x from global memory costs one L2 request
// per nonzero; on MI355X that request rate, not HBM, bounds it.  This layout
// moves the reuse of x into LDS instead:
//   * a row block (<= 32768 rows) is owned by one to eight 1024-thread
//     workgroups (one per column group, below), each of which keeps the
//     block's y partial sums in LDS (128 KiB);
//   * the columns are cut into 8064-wide slices; for every slice in which
//     the block has entries (a "tile"; a slice with more than 32 chunks is
//     cut into several tiles) the workgroup stages x[slice] into LDS with
//     coalesced float4 loads.  kXtDepth tiles ahead of use, their x slices
//     and entry chunks are already in flight in as many register sets (the
//     tile loop is unrolled by kXtDepth so that each set stays in place),
//     which covers the HBM/MALL latency that a one-tile prefetch leaves
//     exposed;
//   * a tile's entries are grouped by row (each row's entries of the slice
//     contiguous; rows of 1-2 entries first, then of 3-16, then longer, so
//     that chunks are alike), packed as (row_in_block << 16 | col_in_slice)
//     plus the value (FP32: a lane's two entries and two values in one
//     16-byte word), and padded to 128-entry chunks; rows of 2-16 entries
//     start on a lane and rows of 3-16 do not cross a chunk boundary.  A
//     wavefront reduces one chunk (2 entries per lane).  A chunk in which no
//     row crosses a lane boundary (flagged by the layout) adds each lane's
//     sums to y_l directly; the others run a segmented scan over the lane
//     tails whose lane exchanges are all DPP moves (no trip through the LDS
//     crossbar), with only the steps the chunk's plan (also from the
//     layout) needs.  Each row's sum is added by one lane, unless the row
//     continues the previous chunk of the tile (the chunk's carry bit): that
//     partial sum goes to the chunk's carry slot.  A tile's up to kXtRounds
//     rounds of chunks run back to back; after the next tile's first
//     barrier one wave sums the carry slots of a tile that has any, per run
//     of equal rows (another DPP scan), and adds each run once.
//   * x is staged once per row block (the x traffic is row blocks x x
//     bytes), so row blocks are as tall as LDS allows; to still give every
//     CU a workgroup, the slices are split into G <= 8 contiguous column
//     groups: workgroup b works on row block b / G and group b % G and
//     writes its partial sums to partial[group][row]; a second kernel adds
//     the G partials to y in group order.
// No atomics: the summation order is fixed by the layout, so results are
// deterministic.  Every workgroup streams the x of its group's slices that
// it touches through LDS, in full cache lines, against one L2 request per
// nonzero for a gather.
//
// Mixed-precision values (the CODEC template argument; FP32 products and
// sums): BF16 hi + lo words (4 bytes per entry, as FP32), FP16 divided by a
// power-of-two row scale or int16 fixed point over a row scale, the scale
// applied once per row when the block's sums leave LDS (2 bytes), or FP8
// E4M3 times one FP32 scale per 128-entry chunk (1 byte).
#include "kcodec.hpp"
#include "kcommon.hpp"
#include "launch.hpp"

namespace clspmv {
namespace dev {

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));

// x[c .. c+3] is read as one vector from min(c, cols - 4) (cols >= 4 is
// required by the host), so the load is unconditional and never leaves x.
// Lanes past the end then hold x shifted by d = c - (cols - 4); x_store4
// shifts the valid elements back (elements at or past cols are never read).
__device__ __forceinline__ f32x4u load_x4(const float* __restrict__ x, int c, int cols) {
    return *reinterpret_cast<const f32x4u*>(x + min(c, cols - 4));
}
__device__ __forceinline__ float4 x_store4(f32x4u v, int c, int cols) {
    const int d = c - (cols - 4);
    float4 o;
    o.x = d <= 0 ? v.x : (d == 1 ? v.y : (d == 2 ? v.z : v.w));
    o.y = d <= 0 ? v.y : (d == 1 ? v.z : v.w);
    o.z = d <= 0 ? v.z : v.w;
    o.w = v.w;
    return o;
}


// Emits a lane's N segment sums (site i: on[i], row[i], s[i]).  A sum is
// added to its row of y_l unless it is padding or the chunk's first row
// continues the previous chunk of the tile (`carry`): that one is stored
// into the chunk's carry slot.  The additions are LDS read-modify-writes
// with all N reads issued before the writes (one LDS round trip); a site
// that adds nothing -- or adds padding, whose row is a dummy slot -- rewrites
// a dummy slot instead, so there is no branch (dummy values are never used).  The rows added to are distinct:
// within one round a row is added to by one lane of one wave only, and the
// sites of one lane end different segments of a row-sorted chunk.
template <int N>
__device__ __forceinline__ void xt_emit(const bool (&on)[N], const int (&row)[N], const float (&s)[N], int first_row,
                                        bool carry, float* y_l, float* dummy, float* cv, int lane) {
    float* p[N];
    float old[N];
    bool fc = false;
    float fv = 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const bool first = on[i] && row[i] == first_row;
        const bool direct = on[i] && !(carry && first);
        p[i] = direct ? &y_l[row[i]] : &dummy[lane];
        fv = first ? s[i] : fv;
        fc = fc || first;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) old[i] = *p[i];
#pragma unroll
    for (int i = 0; i < N; ++i) *p[i] = old[i] + s[i];
    if (carry && fc) *cv = fv;
}

// Lane l receives v of lane l - 1 (lane 0 keeps its own): the DPP
// shift-the-whole-wave-right-by-one mode (control 0x138) of GFX9-family
// parts, a VALU move instead of a trip through the LDS crossbar.
__device__ __forceinline__ int lane_up1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ float lane_up1(float v) { return __int_as_float(lane_up1(__float_as_int(v))); }

// v of another lane by a DPP mode of GFX9-family parts: CTRL 0x110 + d is
// "the lane d below in the same 16-lane row", 0x142 / 0x143 broadcast lane
// 15 / 31 into the next row(s); ROWS selects the rows written (the others,
// and lanes with no source, read 0).
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false); }
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f(float v) { return __int_as_float(dpp_i<CTRL, ROWS>(__float_as_int(v))); }

// One step of a segmented inclusive scan: a lane with no segment start in
// its range so far takes the range below (value added, start flag taken).
__device__ __forceinline__ void seg_step(bool valid, float up_v, int up_f, float& v, int& f) {
    const bool take = valid && !f;
    v = take ? v + up_v : v;
    f = take ? up_f : f;
}

// One step of a segmented inclusive scan over lanes whose keys are sorted:
// a lane adds the partial sum of its DPP source lane when that lane has the
// same key (then so has every lane between).  Sources outside the written
// rows, or outside the wave, read key 0 (no lane's).
template <int CTRL, int ROWS>
__device__ __forceinline__ void xt_scan_step(int key, float& c) {
    const int k = dpp_i<CTRL, ROWS>(key);
    const float u = dpp_f<CTRL, ROWS>(c);
    c = k == key ? c + u : c;
}

static_assert(kXtRounds * kXtWaves == 32, "the carry slots of a tile fill lanes 0..31 of one wave");

// Adds the carry sets of the chunk rounds in `pending` (bit k: round k) to
// y_l, by one whole wave: lane 16 k + i holds slot i of round k (chunk
// order; row -1 marks an unused slot, as do rounds not in `pending`), runs
// of equal rows are summed by a segmented scan in DPP moves (row shifts,
// then lane 15 into lanes 16..31), and the last lane of each run adds the
// run's sum to its row -- one LDS addition per row.
__device__ __forceinline__ void xt_apply_carries(int pending, int (*c_row)[kXtWaves], float (*c_val)[kXtWaves],
                                                 float* y_l, int lane) {
    const int k = (lane >> 4) & 1, i = lane & 15;  // lanes 32..63 mirror 0..31 (unused)
    const bool use = lane < 32 && ((pending >> k) & 1);
    const int r = use ? c_row[k][i] : -1;
    float v = use ? c_val[k][i] : 0.f;
    const int r_prev = lane_up1(r);
    int flag = lane == 0 || r != r_prev;
    const int rl = lane & 15;
    seg_step(rl >= 1, dpp_f<0x111, 0xf>(v), dpp_i<0x111, 0xf>(flag), v, flag);
    seg_step(rl >= 2, dpp_f<0x112, 0xf>(v), dpp_i<0x112, 0xf>(flag), v, flag);
    seg_step(rl >= 4, dpp_f<0x114, 0xf>(v), dpp_i<0x114, 0xf>(flag), v, flag);
    seg_step(rl >= 8, dpp_f<0x118, 0xf>(v), dpp_i<0x118, 0xf>(flag), v, flag);
    seg_step((lane & 16) != 0, dpp_f<0x142, 0xa>(v), dpp_i<0x142, 0xa>(flag), v, flag);
    // the run's last slot: the next lane holds another row (lanes 31.. end)
    const int r_next = __builtin_amdgcn_update_dpp(-1, r, 0x130, 0xf, 0xf, false);  // lane + 1 (wave shift left by one)
    // an LDS add without return: the wave does not wait for y_l's old value
    // (no other wave touches y_l until the next barrier, and the rows are
    // distinct, so the sums are those of a read-modify-write)
    if (lane < 32 && r >= 0 && (lane == 31 || r_next != r)) atomicAdd(&y_l[r], v);
}

// One wavefront reduces one 128-entry chunk (2 entries per lane).  Each
// segment of equal rows is emitted once, by the lane holding its last
// entry: at most two emit sites per lane (the lane's head row when its two
// rows differ, its tail row when the next lane does not go on with it), all
// evaluated without branches.  The chunk's first row goes to the chunk's
// carry slot when its first entry has the carry bit (the row continues the
// previous chunk of the tile), to y_l otherwise.  With `scan_prio` (set
// for layouts whose chunks are mostly fast) a wave on a scan chunk raises
// its issue priority for the chunk: those waves are the ones the next
// barrier waits for.
__device__ __forceinline__ void xt_chunk(u32x2 e, f32x2 v, const float* x_l, float* y_l, float* dummy, int* cr,
                                         float* cv, int wave, int lane, bool tile_carry, bool scan_prio) {
    // products rounded, then added: no multiply-add contraction, so that
    // every value codec (and every compilation) sums the same FP32 numbers
#pragma clang fp contract(off)
    const int ra = (int)(e.x >> 16), rb = (int)(e.y >> 16);
    const float pa = v.x * x_l[e.x & kXtColMask];
    const float pb = v.y * x_l[e.y & kXtColMask];
    const int head_row = ra;
    const bool brk = ra != rb;
    const float head_sum = pa;
    const int tail_row = rb;
    const float tail_sum = brk ? pb : pa + pb;
    const int flags = __builtin_amdgcn_readfirstlane((int)e.x);  // every lane is active here
    if (flags & (int)kXtFastBit) {
        // No row crosses a lane boundary and none continues the previous
        // chunk: each lane adds its entries' sums to their rows directly --
        // entry a alone when the lane's two rows differ, the lane's tail sum
        // to row b (padding, whose row is a dummy slot, adds there).
        if (tile_carry && lane == 0) cr[wave] = -1;  // the slots are read only in tiles with a carry
        float* p1 = brk ? &y_l[head_row] : &dummy[lane];
        float* p2 = &y_l[tail_row];
        const float o1 = *p1, o2 = *p2;
        *p1 = o1 + head_sum;
        *p2 = o2 + tail_sum;
        return;
    }
    if (scan_prio) __builtin_amdgcn_s_setprio(1);
    const int first_row = __builtin_amdgcn_readfirstlane(head_row);
    const bool carry = (flags & (int)kXtCarryBit) != 0;
    if (tile_carry && lane == 0) cr[wave] = carry ? first_row : -1;
    float* slot = &cv[wave];

    // Segmented inclusive scan of the lane tails (row b's sums) in DPP
    // moves, as far as the chunk's scan plan (from the layout) asks: shifts
    // of 1 .. 2^(steps-1) lanes within each 16-lane row cover every run of
    // equal tail rows there, then lane 15 (47) into the next row and lane 31
    // into rows 2 and 3 when a run crosses those lanes.
    const int plan = __builtin_amdgcn_readfirstlane((int)e.y);
    const int steps = ((plan >> 13) & 3) + 1;
    const int key = tail_row + 1;  // lanes with no source read 0: never a row
    float c = tail_sum;
    xt_scan_step<0x111, 0xf>(key, c);
    if (steps >= 2) xt_scan_step<0x112, 0xf>(key, c);
    if (steps >= 3) xt_scan_step<0x114, 0xf>(key, c);
    if (steps >= 4) xt_scan_step<0x118, 0xf>(key, c);
    if (plan & (int)kXtBc15Bit) xt_scan_step<0x142, 0xa>(key, c);
    if (flags & (int)kXtBc31Bit) xt_scan_step<0x143, 0xc>(key, c);
    // a lane's head row continues the previous lane's tail row; its tail row
    // ends here when the next lane's head row is another (past lane 63: -1).
    // The lane moves are made with every lane active: a DPP read from a lane
    // that is off in EXEC does not return that lane's value.
    const int prev_tail = lane_up1(tail_row);
    const int next_head = __builtin_amdgcn_update_dpp(-1, head_row, 0x130, 0xf, 0xf, false);
    const float prev_c = lane_up1(c);
    const bool cont = (lane > 0) & (prev_tail == head_row);
    const bool on[2] = {brk, next_head != tail_row};
    const int row[2] = {head_row, tail_row};
    const float sum[2] = {head_sum + (cont ? prev_c : 0.f), c};
    xt_emit<2>(on, row, sum, first_row, carry, y_l, dummy, slot, lane);
    if (scan_prio) __builtin_amdgcn_s_setprio(0);
}

// Per-lane value storage of one chunk (two entries) for each codec.
template <int CODEC> struct XtVal { typedef f32x2 T; };
template <> struct XtVal<kCodecBf16x2> { typedef u32x2 T; };
template <> struct XtVal<kCodecFp16Row> { typedef unsigned T; };        // entry 0 in the low half
template <> struct XtVal<kCodecI16Row> { typedef unsigned T; };         // entry 0 in the low half
template <> struct XtVal<kCodecFp8Slice> { typedef unsigned short T; };  // entry 0 in the low byte

template <int CODEC>
__device__ __forceinline__ f32x2 xt_decode(typename XtVal<CODEC>::T v, float s) {
    if constexpr (CODEC == kCodecBf16x2) {
        return f32x2{bf16x2_value(v.x), bf16x2_value(v.y)};
    } else if constexpr (CODEC == kCodecFp16Row) {
        return f32x2{f16_value((uint16_t)(v & 0xffffu)), f16_value((uint16_t)(v >> 16))};
    } else if constexpr (CODEC == kCodecI16Row) {
        return f32x2{i16_lo(v), i16_hi(v)};
    } else if constexpr (CODEC == kCodecFp8Slice) {
        return f32x2{s * e4m3_value(v & 0xffu), s * e4m3_value((unsigned)v >> 8)};
    } else {
        (void)s;
        return v;
    }
}

// Registers holding one tile ahead of use: its x slice (unless the tile
// continues the previous tile's slice), this wavefront's chunks, and the
// tile's metadata (wave-uniform), read when the tile was fetched.
template <int CODEC>
struct XtSet {
    f32x4u x[kXtXLoads];
    u32x2 e[kXtRounds];
    typename XtVal<CODEC>::T v[kXtRounds];
    float s[kXtRounds];  // FP8 chunk scales
    int ts, c0, c1;      // slice word and chunk range
};

struct XtArgs {
    const int* blk_row;
    const int* blk_tile;
    const int* tile_slice;  // (slice << 2) | 2 if the tile has a carry | 1 if it continues the previous slice
    const int* tile_chunk;
    const u32x2* idx2;
    const void* val;      // FP8: chunk values, XtVal<CODEC>::T per lane (the other codecs' are in idx2's records)
    const float* vscale;  // FP16: row scales; FP8: chunk scales
    int cols;
    int scan_prio;  // xt_chunk
    const float* x;
    float* y;
};

// Loads are issued unconditionally (indices clamped to valid ones) and each
// register set is refilled in place right after its last use, so that the
// wait before each use covers only the loads issued before it.
template <int CODEC>
__device__ __forceinline__ void xt_fetch_x(XtSet<CODEC>& S, const XtArgs& a, int ts, int tid) {
    S.ts = ts;
    const int base = (ts >> 2) * kXtCols;
#pragma unroll
    for (int q = 0; q < kXtXLoads; ++q)
        S.x[q] = load_x4(a.x, base + min(q * kXtThreads + tid, kXtX4 - 1) * 4, a.cols);
}
struct XtW3 {  // a lane's 12-byte record (loaded as one 12-byte access)
    unsigned w[3];
};
template <int CODEC>
__device__ __forceinline__ void xt_fetch_e(XtSet<CODEC>& S, const XtArgs& a, int c0, int c1, int wave, int lane) {
    typedef typename XtVal<CODEC>::T V;
    S.c0 = c0;
    S.c1 = c1;
#pragma unroll
    for (int k = 0; k < kXtRounds; ++k) {
        // rounds past the tile's chunks load nothing (workgroup-uniform)
        if (k > 0 && c1 - c0 <= k * kXtWaves) break;
        // wave-uniform chunk index (the FP8 scale becomes a scalar load);
        // a wave with no chunk in this chunk round reloads the tile's last chunk
        // (a cache hit) rather than fetching the next tile's entries twice
        const int c = __builtin_amdgcn_readfirstlane(min(c0 + wave + k * kXtWaves, c1 - 1));
        if constexpr (CODEC == 0 || CODEC == kCodecBf16x2) {
            // FP32, BF16 hi+lo: a lane's two entries and two values in one
            // 16-byte record, one load (half the load instructions of two
            // streams)
            const u32x4 q = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(a.idx2) + (size_t)c * 64 + lane);
            S.e[k] = u32x2{q.x, q.y};
            if constexpr (CODEC == 0) S.v[k] = f32x2{__uint_as_float(q.z), __uint_as_float(q.w)};
            else S.v[k] = u32x2{q.z, q.w};
        } else if constexpr (codec_row_scaled(CODEC)) {
            // FP16, int16: two entries and the two 2-byte values, 12 bytes
            const XtW3* r = reinterpret_cast<const XtW3*>(a.idx2) + (size_t)c * 64 + lane;
            const unsigned e0 = __builtin_nontemporal_load(&r->w[0]), e1 = __builtin_nontemporal_load(&r->w[1]);
            S.e[k] = u32x2{e0, e1};
            S.v[k] = __builtin_nontemporal_load(&r->w[2]);
        } else {
            S.e[k] = __builtin_nontemporal_load(a.idx2 + (size_t)c * 64 + lane);
            S.v[k] = __builtin_nontemporal_load(reinterpret_cast<const V*>(a.val) + (size_t)c * 64 + lane);
        }
        if constexpr (CODEC == kCodecFp8Slice) S.s[k] = a.vscale[c];
    }
}

struct XtState {
    int pending;  // bit k: the carry set of chunk round k awaits its addition
};

// Metadata of 64 consecutive tiles of the workgroup, lane l holding tile
// base + l, read back with a wave-uniform lane index: a tile's metadata
// costs no memory round trip (scalar loads of it, waited on before each
// barrier with the LDS operations, were on every tile's critical path);
// the block is reloaded every 64 tiles.
struct XtMeta {
    int base;
    int ts, c0, c1;  // slice word, chunk range of tile base + lane
};
__device__ __forceinline__ void xt_meta_load(XtMeta& M, const XtArgs& a, int base, int t1, int lane) {
    M.base = base;
    const int t = min(base + lane, t1 - 1);
    M.ts = a.tile_slice[t];
    M.c0 = a.tile_chunk[t];
    M.c1 = a.tile_chunk[t + 1];
    // wait here, on the reload path only (vmcnt(0); expcnt and lgkmcnt
    // untouched): otherwise the compiler waits for every load in flight at
    // the reads after the branch, on every tile
    __builtin_amdgcn_s_waitcnt(0x0F70);
}
// Metadata of tile t (wave-uniform, t >= M.base; tiles are asked for in
// increasing order).
__device__ __forceinline__ void xt_meta(XtMeta& M, const XtArgs& a, int t, int t1, int lane, int& ts, int& c0,
                                        int& c1) {
    if (t - M.base >= 64) xt_meta_load(M, a, t, t1, lane);
    const int i = t - M.base;
    ts = __builtin_amdgcn_readlane(M.ts, i);
    c0 = __builtin_amdgcn_readlane(M.c0, i);
    c1 = __builtin_amdgcn_readlane(M.c1, i);
}

// One tile: stage its x slice from set S, refill S with tile t + kXtDepth,
// reduce its chunks.  Steps past the block's last tile (t >= t1) only keep
// the barrier sequence and the carry bookkeeping, so that the unrolled loop
// issues every load on every path.  The chunk rounds of a tile need no
// barrier between them: a row gets a direct addition from at most one
// chunk of the tile (the one its entries start in, unless they start the
// chunk), its other chunks' sums go to their carry slots, and each round
// has its own carry set, added after the next tile's first barrier.
template <int CODEC>
__device__ __forceinline__ void xt_step(XtSet<CODEC>& S, XtMeta& M, const XtArgs& a, int t, int t1, float* y_l,
                                        float4* x_l4, float* dummy, int (*c_row)[kXtWaves],
                                        float (*c_val)[kXtWaves], XtState& st, int tid, int wave, int lane) {
    // the metadata of the tile this step prefetches (the current tile's
    // came with its register set)
    const int tn = min(t + kXtDepth, t1 - 1);
    int ts_n, c0_n, c1_n;
    xt_meta(M, a, tn, t1, lane, ts_n, c0_n, c1_n);
    __syncthreads();  // previous tile done with x_l; its carry slots are written
    if (st.pending && wave == 0) xt_apply_carries(st.pending, c_row, c_val, y_l, lane);
    st.pending = 0;
    const bool live = t < t1;
    const int ts = S.ts;  // tile min(t, t1 - 1)
    if (live && !(ts & 1)) {
        const int base = (ts >> 2) * kXtCols;
        if (base + kXtCols <= a.cols) {
#pragma unroll
            for (int q = 0; q < kXtXLoads; ++q) {
                const int i = q * kXtThreads + tid;
                const f32x4u v = S.x[q];
                if (kXtX4 % kXtThreads == 0 || i < kXtX4) x_l4[i] = make_float4(v.x, v.y, v.z, v.w);
            }
        } else {  // the slice holding the end of x
#pragma unroll
            for (int q = 0; q < kXtXLoads; ++q) {
                const int i = q * kXtThreads + tid;
                if (kXtX4 % kXtThreads == 0 || i < kXtX4) x_l4[i] = x_store4(S.x[q], base + i * 4, a.cols);
            }
        }
    }
    const int c0 = S.c0, c1 = live ? S.c1 : c0;
    xt_fetch_x(S, a, ts_n, tid);
    __syncthreads();  // x_l holds the slice; carries applied
    const float* x_l = reinterpret_cast<const float*>(x_l4);
#pragma unroll
    for (int k = 0; k < kXtRounds; ++k) {
        const int cb = c0 + k * kXtWaves;
        if (cb < c1) {
            if (cb + wave < c1)
                xt_chunk(S.e[k], xt_decode<CODEC>(S.v[k], S.s[k]), x_l, y_l, dummy, c_row[k], c_val[k], wave, lane,
                         (ts & 2) != 0, a.scan_prio != 0);
            else if ((ts & 2) && lane == 0)
                c_row[k][wave] = -1;
            if (ts & 2) st.pending |= 1 << k;  // the tile has a carry
        }
    }
    xt_fetch_e(S, a, c0_n, c1_n, wave, lane);
}

template <int CODEC>
__global__ __launch_bounds__(kXtThreads) void xtile_kernel(
    const int* __restrict__ blk_row, const int* __restrict__ blk_tile, const int* __restrict__ tile_slice,
    const int* __restrict__ tile_chunk, const u32x2* __restrict__ idx2, const void* __restrict__ val,
    const float* __restrict__ vscale, int cols, const float* __restrict__ x, float* __restrict__ y,
    float* __restrict__ partial, int groups, int rows, int scan_prio) {
    // the metadata is read-only and unaliased, so its wave-uniform reads
    // become scalar loads that do not wait on the vector-load queue
    const XtArgs a{blk_row, blk_tile, tile_slice, tile_chunk, idx2, val, vscale, cols, scan_prio, x, y};
    __shared__ float y_l[kXtRows + 64];  // + one dummy slot per lane (the first is the padding row)
    __shared__ float4 x_l4[kXtX4];
    __shared__ int c_row[kXtRounds][kXtWaves];
    __shared__ float c_val[kXtRounds][kXtWaves];
    float* dummy = y_l + kXtRows;  // where a lane's non-emits add
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int b = blockIdx.x;
    const int rb = b / groups;  // row block; column group b % groups
    const int r0 = a.blk_row[rb], nr = a.blk_row[rb + 1] - r0;
    const int t0 = a.blk_tile[b], t1 = a.blk_tile[b + 1];
    float* out = groups > 1 ? partial + (size_t)(b % groups) * rows + r0 : nullptr;
    if (t0 == t1) {  // no entries in this row block and group (uniform per workgroup)
        if (out)
            for (int r = tid; r < nr; r += kXtThreads) out[r] = 0.f;
        return;
    }
    for (int r = tid; r < nr; r += kXtThreads) y_l[r] = 0.f;

    // kXtDepth register sets, indexed only by constants (the loops over them
    // are unrolled), so they stay in registers
    XtSet<CODEC> S[kXtDepth];
    XtMeta M;
    xt_meta_load(M, a, t0, t1, lane);
#pragma unroll
    for (int u = 0; u < kXtDepth; ++u) {
        int ts, c0, c1;
        xt_meta(M, a, min(t0 + u, t1 - 1), t1, lane, ts, c0, c1);
        xt_fetch_x(S[u], a, ts, tid);
        xt_fetch_e(S[u], a, c0, c1, wave, lane);
    }
    XtState st{0};
    for (int t = t0; t < t1; t += kXtDepth) {
#pragma unroll
        for (int u = 0; u < kXtDepth; ++u)
            xt_step(S[u], M, a, t + u, t1, y_l, x_l4, dummy, c_row, c_val, st, tid, wave, lane);
    }
    __syncthreads();
    if (st.pending && wave == 0) xt_apply_carries(st.pending, c_row, c_val, y_l, lane);
    __syncthreads();
    if constexpr (codec_row_scaled(CODEC))
        for (int r = tid; r < nr; r += kXtThreads) y_l[r] *= a.vscale[r0 + r];
    if (out) {
        for (int r = tid; r < nr; r += kXtThreads) out[r] = y_l[r];
    } else {
        for (int r = tid; r < nr; r += kXtThreads) a.y[r0 + r] += y_l[r];
    }
}

// y[r] += partial[0][r] + partial[1][r] + ... in group order, in FP64 (the
// cross-workgroup part of a row's sum).
__global__ __launch_bounds__(256) void xtile_reduce_kernel(const float* __restrict__ partial, int groups, int rows,
                                                           float* __restrict__ y) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= rows) return;
    double s = partial[r];
    for (int g = 1; g < groups; ++g) s += (double)partial[(size_t)g * rows + r];
    y[r] = (float)((double)y[r] + s);
}

// The same sums, four rows per thread in 16-byte loads and stores (rows a
// multiple of 4, y 16-byte aligned; the partial rows are then too).
__global__ __launch_bounds__(256) void xtile_reduce4_kernel(const float* __restrict__ partial, int groups, int rows,
                                                            float* __restrict__ y) {
    const int r = (blockIdx.x * 256 + threadIdx.x) * 4;
    if (r >= rows) return;
    const float4 p0 = *reinterpret_cast<const float4*>(partial + r);
    double s0 = p0.x, s1 = p0.y, s2 = p0.z, s3 = p0.w;
    for (int g = 1; g < groups; ++g) {
        const float4 p = *reinterpret_cast<const float4*>(partial + (size_t)g * rows + r);
        s0 += (double)p.x, s1 += (double)p.y, s2 += (double)p.z, s3 += (double)p.w;
    }
    float4* yv = reinterpret_cast<float4*>(y + r);
    const float4 o = *yv;
    *yv = make_float4((float)((double)o.x + s0), (float)((double)o.y + s1), (float)((double)o.z + s2),
                      (float)((double)o.w + s3));
}

}  // namespace dev

int launch_xtile(const DevXtile& m, const float* x, float* y, hipStream_t st) {
    if (m.nblocks == 0) return 0;
#define CLSPMV_XTILE(C)                                                                                          \
    hipLaunchKernelGGL(dev::xtile_kernel<C>, dim3(m.nblocks), dim3(kXtThreads), 0, st, m.blk_row, m.blk_tile, \
                       m.tile_slice, m.tile_chunk, reinterpret_cast<const dev::u32x2*>(m.idx), m.val, m.vscale,   \
                       m.cols, x, y, m.partial, m.groups, m.rows, m.scan_prio)
    switch (m.codec) {
        case kCodecBf16x2: CLSPMV_XTILE(kCodecBf16x2); break;
        case kCodecFp16Row: CLSPMV_XTILE(kCodecFp16Row); break;
        case kCodecI16Row: CLSPMV_XTILE(kCodecI16Row); break;
        case kCodecFp8Slice: CLSPMV_XTILE(kCodecFp8Slice); break;
        default: CLSPMV_XTILE(0); break;
    }
#undef CLSPMV_XTILE
    int e = (int)hipGetLastError();
    if (e || m.groups == 1) return e;
    if (m.rows % 4 == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0)
        hipLaunchKernelGGL(dev::xtile_reduce4_kernel, dim3((m.rows / 4 + 255) / 256), dim3(256), 0, st, m.partial,
                           m.groups, m.rows, y);
    else
        hipLaunchKernelGGL(dev::xtile_reduce_kernel, dim3((m.rows + 255) / 256), dim3(256), 0, st, m.partial,
                           m.groups, m.rows, y);
    return (int)hipGetLastError();
}

}  // namespace clspmv
