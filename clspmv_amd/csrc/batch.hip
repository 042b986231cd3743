// Batched multi-right-hand-side SpMV: host side (device pools, work-item
// planning, hipGraph capture) of k_spmm.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "clspmv.h"
#include "clspmv_batch.h"
#include "host/internal.hpp"
#include "kernels/launch.hpp"
#include "runtime.hpp"

using namespace clspmv;

struct clspmv_batch_s {
    clspmv_ctx_t ctx = nullptr;
    DevBuf rp, col, val, x, y, mats, items, longs, carry;
    std::vector<SpmmMat> host_mats;
    std::vector<size_t> x_off, y_off;  // float offsets of X_i, Y_i in the pools
    std::vector<int> rows, cols, ks;
    DevSpmm dev{};
    long long nnz = 0, nnz_rhs = 0, alg_bytes = 0;
    int n_block_mats = 0;  // BSR matrices (BLOCK items on the matrix cores)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~clspmv_batch_s() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (DevBuf* b : {&rp, &col, &val, &x, &y, &mats, &items, &longs, &carry}) b->release();
    }
};

namespace {

size_t align16(size_t n) { return (n + 15) / 16 * 16; }

int check_csr(const clspmv_csr_matrix* m) {
    const int rows = m->matinfo.height, cols = m->matinfo.width;
    if (rows < 0 || cols < 0 || !m->csr_row_ptr || m->csr_row_ptr[0] != 0) return CLSPMV_E_INVALID;
    for (int r = 0; r < rows; ++r)
        if (m->csr_row_ptr[r + 1] < m->csr_row_ptr[r]) return CLSPMV_E_INVALID;
    const int nnz = m->csr_row_ptr[rows];
    for (int k = 0; k < nnz; ++k)
        if (m->csr_col_id[k] < 0 || m->csr_col_id[k] >= cols) {
            set_error("batch: column index out of range");
            return CLSPMV_E_INVALID;
        }
    return CLSPMV_OK;
}

// Items of matrix i: STREAM blocks of consecutive rows within the tile's
// caps, LONG chunks for rows above the threshold.
void plan(int i, const clspmv_csr_matrix* m, const clspmv_spmm_tile& t, std::vector<int4>& items,
          std::vector<int4>& longs, int& slots) {
    const int* rp = m->csr_row_ptr;
    const int rows = m->matinfo.height;
    int r = 0;
    while (r < rows) {
        const int len = rp[r + 1] - rp[r];
        if (len > t.long_threshold) {
            const int chunks = (len + kSpmmLongChunk - 1) / kSpmmLongChunk;
            longs.push_back(make_int4(i, r, slots, chunks));
            for (int c = 0; c < chunks; ++c) items.push_back(make_int4(i, r, -(slots + c + 1), c));
            slots += chunks;
            ++r;
            continue;
        }
        const int r0 = r;
        int n = 0;
        while (r < rows && r - r0 < t.rows_cap) {
            const int l = rp[r + 1] - rp[r];
            if (l > t.long_threshold || (r > r0 && n + l > t.nnz_cap)) break;
            n += l;
            ++r;
        }
        items.push_back(make_int4(i, r0, r, 0));
    }
}

// BLOCK items of a BSR matrix: 16 / (k / 4) block rows each.
void plan_blocks(int i, const clspmv_bsr_matrix* m, int k, std::vector<int4>& items) {
    const int per = 16 / (k / 4), brows = m->bsr_row_num;
    for (int b = 0; b < brows; b += per) items.push_back(make_int4(i, b, std::min(brows, b + per), 0));
}

int check_bsr(const clspmv_bsr_matrix* m, int k) {
    const unsigned b = m->bsr_bwidth;
    const int rows = m->matinfo.height, cols = m->matinfo.width, brows = m->bsr_row_num, nb = m->bsr_block_num;
    if (b != m->bsr_bheight || b < 2 || b > 4 || (k != 4 && k != 16) || rows <= 0 || cols <= 0 ||
        brows != (rows + (int)b - 1) / (int)b || !m->bsr_row_ptr || m->bsr_row_ptr[0] != 0 ||
        m->bsr_row_ptr[brows] != nb || m->bsr_block_stride < nb) {
        set_error("batch: BSR matrix needs square 2x2..4x4 blocks, k 4 or 16 and a consistent layout");
        return CLSPMV_E_INVALID;
    }
    for (int r = 0; r < brows; ++r)
        if (m->bsr_row_ptr[r + 1] < m->bsr_row_ptr[r]) return CLSPMV_E_INVALID;
    for (int j = 0; j < nb; ++j)
        if (m->bsr_col_id[j] < 0 || m->bsr_col_id[j] >= cols) {
            set_error("batch: BSR block column out of range");
            return CLSPMV_E_INVALID;
        }
    return CLSPMV_OK;
}

bool tile_ok(const clspmv_spmm_tile& t) {
    return t.rows_cap >= 1 && t.rows_cap <= 4096 && t.nnz_cap >= 1 && t.nnz_cap <= kSpmmStageNnz &&
           t.long_threshold >= 1 && t.long_threshold <= t.nnz_cap;
}

}  // namespace

extern "C" {

clspmv_spmm_tile clspmv_spmm_default_tile(int k) {
    clspmv_spmm_tile t;
    t.rows_cap = k >= 16 ? 32 : 64;
    t.nnz_cap = kSpmmStageNnz;
    t.long_threshold = 256;
    return t;
}

int clspmv_batch_create(clspmv_ctx_t ctx, int n, const clspmv_csr_matrix* const* mats, const int* ks,
                        const clspmv_spmm_tile* tiles, clspmv_batch_t* out) {
    if (!ctx || n <= 0 || !mats || !ks || !out) return CLSPMV_E_INVALID;
    std::vector<clspmv_batch_entry> e((size_t)n);
    for (int i = 0; i < n; ++i) e[(size_t)i] = clspmv_batch_entry{mats[i], nullptr, ks[i]};
    return clspmv_batch_create_entries(ctx, n, e.data(), tiles, out);
}

int clspmv_batch_create_entries(clspmv_ctx_t ctx, int n, const clspmv_batch_entry* ent,
                                const clspmv_spmm_tile* tiles, clspmv_batch_t* out) {
    if (!ctx || n <= 0 || !ent || !out) return CLSPMV_E_INVALID;
    *out = nullptr;
    auto* b = new (std::nothrow) clspmv_batch_s();
    if (!b) return CLSPMV_E_NOMEM;
    b->ctx = ctx;
    size_t nrp = 0, nnz = 0, nx = 0, ny = 0;
    std::vector<size_t> rp_off, nz_off;
    for (int i = 0; i < n; ++i) {
        const clspmv_csr_matrix* m = ent[i].csr;
        const clspmv_bsr_matrix* bm = ent[i].bsr;
        const int k = ent[i].k;
        int rc = (!m == !bm || (k != 1 && k != 4 && k != 16)) ? CLSPMV_E_INVALID
                 : m                                          ? check_csr(m)
                                                              : check_bsr(bm, k);
        if (!rc && tiles && m && !tile_ok(tiles[i])) rc = CLSPMV_E_INVALID;
        if (rc) {
            set_error("batch: matrix %d invalid (one of csr / bsr; k 1, 4 or 16, BSR k 4 or 16; tile caps within "
                      "limits)", i);
            delete b;
            return rc;
        }
        const int rows = m ? m->matinfo.height : bm->matinfo.height;
        const int cols = m ? m->matinfo.width : bm->matinfo.width;
        rp_off.push_back(nrp);
        nz_off.push_back(nnz);
        b->x_off.push_back(nx);
        b->y_off.push_back(ny);
        if (m) {
            nrp += align16((size_t)rows + 1);
            nnz += align16((size_t)m->csr_row_ptr[rows]);
        } else {
            const size_t bs = (size_t)bm->bsr_bwidth * bm->bsr_bheight;
            nrp += align16((size_t)bm->bsr_row_num + 1);
            nnz += align16((size_t)bm->bsr_block_num * bs);  // values; block columns use the first nb slots
        }
        nx += align16((size_t)cols * k);
        ny += align16((size_t)rows * k);
        b->rows.push_back(rows);
        b->cols.push_back(cols);
        b->ks.push_back(k);
    }
    int rc = b->rp.alloc(std::max<size_t>(nrp, 16) * 4);
    if (!rc) rc = b->col.alloc(std::max<size_t>(nnz, 16) * 4);
    if (!rc) rc = b->val.alloc(std::max<size_t>(nnz, 16) * 4);
    if (!rc) rc = b->x.alloc(std::max<size_t>(nx, 16) * 4);
    if (!rc) rc = b->y.alloc(std::max<size_t>(ny, 16) * 4);
    hipStream_t st = ctx->stream();
    if (!rc) rc = hip_status(hipMemsetAsync(b->x.ptr, 0, b->x.size, st), "hipMemset");
    if (!rc) rc = hip_status(hipMemsetAsync(b->y.ptr, 0, b->y.size, st), "hipMemset");
    std::vector<int4> items, longs;
    std::vector<std::vector<float>> staged;  // host copies alive until the uploads finish
    int slots = 0;
    for (int i = 0; i < n && !rc; ++i) {
        const clspmv_csr_matrix* m = ent[i].csr;
        const clspmv_bsr_matrix* bm = ent[i].bsr;
        const int k = ent[i].k;
        int* drp = (int*)b->rp.ptr + rp_off[(size_t)i];
        int* dcol = (int*)b->col.ptr + nz_off[(size_t)i];
        float* dval = (float*)b->val.ptr + nz_off[(size_t)i];
        SpmmMat d{};
        d.row_ptr = drp;
        d.col = dcol;
        d.val = dval;
        d.x = (const float*)b->x.ptr + b->x_off[(size_t)i];
        d.y = (float*)b->y.ptr + b->y_off[(size_t)i];
        d.rows = b->rows[(size_t)i];
        d.cols = b->cols[(size_t)i];
        d.k = k;
        if (m) {
            const size_t rows = (size_t)m->matinfo.height, z = (size_t)m->csr_row_ptr[rows];
            rc = hip_status(hipMemcpyAsync(drp, m->csr_row_ptr, (rows + 1) * 4, hipMemcpyHostToDevice, st), "h2d");
            if (!rc && z) rc = hip_status(hipMemcpyAsync(dcol, m->csr_col_id, z * 4, hipMemcpyHostToDevice, st), "h2d");
            if (!rc && z) rc = hip_status(hipMemcpyAsync(dval, m->csr_data, z * 4, hipMemcpyHostToDevice, st), "h2d");
            plan(i, m, tiles ? tiles[i] : clspmv_spmm_default_tile(k), items, longs, slots);
            b->nnz += (long long)z;
            b->nnz_rhs += (long long)z * k;
            b->alg_bytes += 8LL * (long long)z + 4LL * ((long long)rows + 1) +
                            (long long)k * (4LL * m->matinfo.width + 8LL * (long long)rows);
        } else {
            // blocks row-major inside each block (the BSR struct stores
            // element e of block j at data[e * stride + j])
            const int bd = (int)bm->bsr_bwidth, brows = bm->bsr_row_num, nb = bm->bsr_block_num;
            const size_t bs = (size_t)bd * bd;
            staged.emplace_back((size_t)nb * bs);
            std::vector<float>& v = staged.back();
            for (int j = 0; j < nb; ++j)
                for (size_t e = 0; e < bs; ++e) v[(size_t)j * bs + e] = bm->bsr_data[e * bm->bsr_block_stride + j];
            rc = hip_status(hipMemcpyAsync(drp, bm->bsr_row_ptr, ((size_t)brows + 1) * 4, hipMemcpyHostToDevice, st),
                            "h2d");
            if (!rc && nb)
                rc = hip_status(hipMemcpyAsync(dcol, bm->bsr_col_id, (size_t)nb * 4, hipMemcpyHostToDevice, st), "h2d");
            if (!rc && nb) rc = hip_status(hipMemcpyAsync(dval, v.data(), v.size() * 4, hipMemcpyHostToDevice, st), "h2d");
            d.bdim = bd;
            plan_blocks(i, bm, k, items);
            const long long true_nnz = bm->matinfo.nnz;
            b->nnz += true_nnz;
            b->nnz_rhs += true_nnz * k;
            b->alg_bytes += (4LL * (long long)bs + 4LL) * nb + 4LL * ((long long)brows + 1) +
                            (long long)k * (4LL * bm->matinfo.width + 8LL * (long long)bm->matinfo.height);
            ++b->n_block_mats;
        }
        b->host_mats.push_back(d);
    }
    if (!rc) rc = hip_status(hipStreamSynchronize(st), "batch upload");
    staged.clear();
    if (!rc) rc = b->mats.upload(ctx, b->host_mats.data(), b->host_mats.size() * sizeof(SpmmMat));
    if (!rc) rc = b->items.upload(ctx, items.data(), std::max<size_t>(1, items.size()) * sizeof(int4));
    if (!rc) rc = b->longs.upload(ctx, longs.data(), std::max<size_t>(1, longs.size()) * sizeof(int4));
    if (!rc) rc = b->carry.alloc((size_t)std::max(1, slots) * 16 * 4);
    if (rc) {
        delete b;
        return rc;
    }
    b->dev.mats = (const SpmmMat*)b->mats.ptr;
    b->dev.items = (const int4*)b->items.ptr;
    b->dev.longs = (const int4*)b->longs.ptr;
    b->dev.carry = (float*)b->carry.ptr;
    b->dev.n_items = (int)items.size();
    b->dev.n_long = (int)longs.size();
    // persistent grid: at most 8 four-wave workgroups per CU
    b->dev.grid = std::max(1, std::min((b->dev.n_items + 3) / 4, 256 * 8));
    *out = b;
    return CLSPMV_OK;
}

int clspmv_batch_free(clspmv_batch_t b) {
    if (!b) return CLSPMV_E_INVALID;
    (void)hipStreamSynchronize(b->ctx->stream());
    delete b;
    return CLSPMV_OK;
}

int clspmv_batch_size(clspmv_batch_t b) { return b ? (int)b->ks.size() : 0; }

int clspmv_batch_buffers(clspmv_batch_t b, int i, void** d_x, void** d_y) {
    if (!b || i < 0 || i >= (int)b->ks.size()) return CLSPMV_E_INVALID;
    if (d_x) *d_x = (float*)b->x.ptr + b->x_off[(size_t)i];
    if (d_y) *d_y = (float*)b->y.ptr + b->y_off[(size_t)i];
    return CLSPMV_OK;
}

int clspmv_batch_set_x(clspmv_batch_t b, int i, const float* x) {
    if (!b || !x || i < 0 || i >= (int)b->ks.size()) return CLSPMV_E_INVALID;
    return clspmv_memcpy_h2d(b->ctx, (float*)b->x.ptr + b->x_off[(size_t)i], x,
                             (size_t)b->cols[(size_t)i] * b->ks[(size_t)i] * 4);
}
int clspmv_batch_set_y(clspmv_batch_t b, int i, const float* y) {
    if (!b || !y || i < 0 || i >= (int)b->ks.size()) return CLSPMV_E_INVALID;
    return clspmv_memcpy_h2d(b->ctx, (float*)b->y.ptr + b->y_off[(size_t)i], y,
                             (size_t)b->rows[(size_t)i] * b->ks[(size_t)i] * 4);
}
int clspmv_batch_get_y(clspmv_batch_t b, int i, float* y) {
    if (!b || !y || i < 0 || i >= (int)b->ks.size()) return CLSPMV_E_INVALID;
    return clspmv_memcpy_d2h(b->ctx, y, (float*)b->y.ptr + b->y_off[(size_t)i],
                             (size_t)b->rows[(size_t)i] * b->ks[(size_t)i] * 4);
}

int clspmv_batch_run(clspmv_batch_t b, int ntimes) {
    if (!b || ntimes < 0) return CLSPMV_E_INVALID;
    hipStream_t st = b->ctx->stream();
    for (int i = 0; i < ntimes; ++i) {
        if (b->exec) {
            int rc = hip_status(hipGraphLaunch(b->exec, st), "hipGraphLaunch");
            if (rc) return rc;
        } else {
            int e = launch_spmm(b->dev, st);
            if (e) return hip_status((hipError_t)e, "batch launch");
        }
    }
    return CLSPMV_OK;
}

int clspmv_batch_capture(clspmv_batch_t b) {
    if (!b) return CLSPMV_E_INVALID;
    if (b->exec) return CLSPMV_OK;
    hipStream_t st = b->ctx->stream();
    int rc = hip_status(hipStreamSynchronize(st), "sync");
    if (!rc) rc = hip_status(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    if (rc) return rc;
    const int e = launch_spmm(b->dev, st);
    hipGraph_t g = nullptr;
    rc = hip_status(hipStreamEndCapture(st, &g), "hipStreamEndCapture");
    if (!rc && e) rc = hip_status((hipError_t)e, "batch launch under capture");
    if (!rc) rc = hip_status(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        b->exec = nullptr;
        return rc;
    }
    b->graph = g;
    return CLSPMV_OK;
}

int clspmv_batch_stats(clspmv_batch_t b, long long* nnz, long long* nnz_rhs, long long* alg_bytes, int* items,
                       int* grid) {
    if (!b) return CLSPMV_E_INVALID;
    if (nnz) *nnz = b->nnz;
    if (nnz_rhs) *nnz_rhs = b->nnz_rhs;
    if (alg_bytes) *alg_bytes = b->alg_bytes;
    if (items) *items = b->dev.n_items;
    if (grid) *grid = b->dev.grid;
    return CLSPMV_OK;
}

}  // extern "C"
