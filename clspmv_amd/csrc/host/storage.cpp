// Host-side sparse data model: the format conversions of the reference's
// A/include/matrix_storage.h, written for this engine.
//
// Every conversion produces arrays that are bit-identical to the reference's
// for the same sorted COO input (including its padding rules); differences are
// limited to cases where the reference has undefined behaviour (documented at
// each function) and where it would exit/assert (we return a status code).
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "clspmv.h"
#include "internal.hpp"

namespace clspmv {

thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

// Row pointer of a row-major sorted COO: rp[r] = #entries with row < r.
// Same values as the reference's while-loop construction
// (matrix_storage.h:1288-1306) for sorted input with rows in [0, height).
int row_ptr_of(const clspmv_coo_matrix* m, std::vector<int>& rp) {
    const int h = m->matinfo.height;
    const int nnz = m->matinfo.nnz;
    rp.assign((size_t)h + 1, 0);
    for (int k = 0; k < nnz; ++k) {
        int r = m->coo_row_id[k];
        if (r < 0 || r >= h) {
            set_error("row index %d out of range [0,%d) at entry %d", r, h, k);
            return CLSPMV_E_INVALID;
        }
        rp[(size_t)r + 1]++;
    }
    for (int r = 0; r < h; ++r) rp[(size_t)r + 1] += rp[r];
    return CLSPMV_OK;
}

int check_cols(const clspmv_coo_matrix* m) {
    const int w = m->matinfo.width;
    for (int k = 0; k < m->matinfo.nnz; ++k) {
        int c = m->coo_col_id[k];
        if (c < 0 || c >= w) {
            set_error("column index %d out of range [0,%d) at entry %d", c, w, k);
            return CLSPMV_E_INVALID;
        }
    }
    return CLSPMV_OK;
}

int ensure_sorted(clspmv_coo_matrix* m) {
    if (clspmv_if_sorted_coo(m)) return CLSPMV_OK;
    return clspmv_sort_coo(m);
}

template <class T>
static T* xmalloc(size_t n) {
    return static_cast<T*>(std::malloc(n ? n * sizeof(T) : 1));
}
template <class T>
static T* xcalloc(size_t n) {
    return static_cast<T*>(std::calloc(n ? n : 1, sizeof(T)));
}

// Block extraction shared by the blocked formats (the per-block-row loop of
// matrix_storage.h:1646-1712, repeated in coo2b4csr/coo2b4ell/coo2sbell).
// The entries of a block row are ordered by (col, row) with std::sort and the
// same comparator as the reference (compareCol, :1532-1541), so that even
// duplicate (row, col) entries land in the same order; consecutive entries
// with the same block column (col - col % bw) form one dense bh x bw block,
// stored row-major in `data`.
struct BlockRows {
    std::vector<int> rowptr;   // size brows + 1
    std::vector<int> bcol;     // first column of each block
    std::vector<float> data;   // bh*bw per block, row-major inside the block
};

struct Elem {
    int row, col;
    float val;
};

static int extract_blocks(const clspmv_coo_matrix* m, const std::vector<int>& rp,
                          unsigned bw, unsigned bh, BlockRows& out) {
    const int h = m->matinfo.height;
    const int brows = (h + (int)bh - 1) / (int)bh;
    const unsigned bs = bw * bh;
    out.rowptr.assign((size_t)brows + 1, 0);
    int nthreads = omp_threads();
    std::vector<std::vector<int>> tcol(nthreads);
    std::vector<std::vector<float>> tdat(nthreads);
    std::vector<int> tcount((size_t)brows, 0);
    // Contiguous chunks of block rows per thread keep the output order.
    std::vector<int> chunk_begin(nthreads + 1);
    for (int t = 0; t <= nthreads; ++t) chunk_begin[t] = (int)((int64_t)brows * t / nthreads);
#pragma omp parallel num_threads(nthreads)
    {
        int t = omp_tid();
        std::vector<Elem> el;
        for (int br = chunk_begin[t]; br < chunk_begin[t + 1]; ++br) {
            int r0 = br * (int)bh;
            int r1 = std::min(h, r0 + (int)bh);
            int s = rp[r0], e = rp[r1];
            if (e <= s) continue;
            el.resize((size_t)(e - s));
            for (int k = s; k < e; ++k) el[k - s] = {m->coo_row_id[k], m->coo_col_id[k], m->coo_data[k]};
            std::sort(el.begin(), el.end(), [](const Elem& a, const Elem& b) {
                if (a.col != b.col) return a.col < b.col;
                return a.row < b.row;
            });
            int nb = 0;
            size_t i = 0;
            while (i < el.size()) {
                int bc = el[i].col - el[i].col % (int)bw;
                size_t base = tdat[t].size();
                tdat[t].resize(base + bs, 0.0f);
                tcol[t].push_back(bc);
                while (i < el.size() && el[i].col - el[i].col % (int)bw == bc) {
                    unsigned ir = (unsigned)(el[i].row % (int)bh);
                    unsigned ic = (unsigned)(el[i].col - bc);
                    tdat[t][base + ir * bw + ic] = el[i].val;
                    ++i;
                }
                ++nb;
            }
            tcount[br] = nb;
        }
    }
    for (int br = 0; br < brows; ++br) out.rowptr[(size_t)br + 1] = out.rowptr[br] + tcount[br];
    size_t nblocks = (size_t)out.rowptr[brows];
    out.bcol.resize(nblocks);
    out.data.resize(nblocks * bs);
    size_t cpos = 0;
    for (int t = 0; t < nthreads; ++t) {
        std::copy(tcol[t].begin(), tcol[t].end(), out.bcol.begin() + cpos);
        std::copy(tdat[t].begin(), tdat[t].end(), out.data.begin() + cpos * bs);
        cpos += tcol[t].size();
    }
    return CLSPMV_OK;
}

}  // namespace clspmv

using namespace clspmv;

extern "C" {

const char* clspmv_last_error(void) { return g_last_error; }

void clspmv_init_coo(clspmv_coo_matrix* m) { std::memset(m, 0, sizeof(*m)); }

#define FREE_FIELD(p) \
    do {              \
        std::free(p); \
        (p) = nullptr; \
    } while (0)

void clspmv_free_coo(clspmv_coo_matrix* m) {
    FREE_FIELD(m->coo_row_id);
    FREE_FIELD(m->coo_col_id);
    FREE_FIELD(m->coo_data);
}
void clspmv_free_csr(clspmv_csr_matrix* m) {
    FREE_FIELD(m->csr_row_ptr);
    FREE_FIELD(m->csr_col_id);
    FREE_FIELD(m->csr_data);
}
void clspmv_free_dia(clspmv_dia_matrix* m) {
    FREE_FIELD(m->dia_offsets);
    FREE_FIELD(m->dia_data);
}
void clspmv_free_bdia(clspmv_bdia_matrix* m) {
    FREE_FIELD(m->bdia_offsets);
    FREE_FIELD(m->bdia_bptr);
    FREE_FIELD(m->bdia_data);
}
void clspmv_free_ell(clspmv_ell_matrix* m) {
    FREE_FIELD(m->ell_col_id);
    FREE_FIELD(m->ell_data);
}
void clspmv_free_sell(clspmv_sell_matrix* m) {
    FREE_FIELD(m->sell_slice_ptr);
    FREE_FIELD(m->sell_col_id);
    FREE_FIELD(m->sell_data);
}
void clspmv_free_b4csr(clspmv_b4csr_matrix* m) {
    FREE_FIELD(m->b4csr_row_ptr);
    FREE_FIELD(m->b4csr_col_id);
    FREE_FIELD(m->b4csr_data);
}
void clspmv_free_b4ell(clspmv_b4ell_matrix* m) {
    FREE_FIELD(m->b4ell_col_id);
    FREE_FIELD(m->b4ell_data);
}
void clspmv_free_sbell(clspmv_sbell_matrix* m) {
    FREE_FIELD(m->sbell_slice_ptr);
    FREE_FIELD(m->sbell_col_id);
    FREE_FIELD(m->sbell_data);
}
void clspmv_free_bsr(clspmv_bsr_matrix* m) {
    FREE_FIELD(m->bsr_row_ptr);
    FREE_FIELD(m->bsr_col_id);
    FREE_FIELD(m->bsr_data);
}

int clspmv_alloc_coo(clspmv_coo_matrix* m, int width, int height, int nnz) {
    if (width < 0 || height < 0 || nnz < 0) return CLSPMV_E_INVALID;
    m->matinfo.width = width;
    m->matinfo.height = height;
    m->matinfo.nnz = nnz;
    m->coo_row_id = xmalloc<int>((size_t)nnz);
    m->coo_col_id = xmalloc<int>((size_t)nnz);
    m->coo_data = xmalloc<float>((size_t)nnz);
    if (!m->coo_row_id || !m->coo_col_id || !m->coo_data) {
        clspmv_free_coo(m);
        return CLSPMV_E_NOMEM;
    }
    return CLSPMV_OK;
}

int clspmv_aligned_length(int length, int alignment) {
    if (alignment <= 0) return length;
    if (length % alignment == 0) return length;
    return length + alignment - length % alignment;
}

int clspmv_if_sorted_coo(const clspmv_coo_matrix* m) {
    const int nnz = m->matinfo.nnz;
    for (int i = 0; i + 1 < nnz; ++i) {
        int r0 = m->coo_row_id[i], r1 = m->coo_row_id[i + 1];
        if (r0 > r1 || (r0 == r1 && m->coo_col_id[i] > m->coo_col_id[i + 1])) return 0;
    }
    return 1;
}

// Non-recursive quicksort with an explicit stack of at most MAX_LEVELS
// segments, pivot = middle element moved to the segment start; this is the
// algorithm of sort_coo (matrix_storage.h:744-815), kept step for step so that
// entries with equal (row, col) end in the same order as in the reference.
int clspmv_sort_coo(clspmv_coo_matrix* m) {
    int* row = m->coo_row_id;
    int* col = m->coo_col_id;
    float* val = m->coo_data;
    std::vector<int> lo(CLSPMV_MAX_LEVELS), hi(CLSPMV_MAX_LEVELS);
    int depth = 0;
    lo[0] = 0;
    hi[0] = m->matinfo.nnz;
    auto greater = [&](int k, int pr, int pc) { return row[k] > pr || (row[k] == pr && col[k] > pc); };
    auto less = [&](int k, int pr, int pc) { return row[k] < pr || (row[k] == pr && col[k] < pc); };
    while (depth >= 0) {
        int L = lo[depth];
        int R = hi[depth] - 1;
        if (L >= R) {
            --depth;
            continue;
        }
        int mid = (L + R) / 2;
        int pr = row[mid], pc = col[mid];
        float pv = val[mid];
        row[mid] = row[L];
        col[mid] = col[L];
        val[mid] = val[L];
        if (depth == CLSPMV_MAX_LEVELS - 1) {
            // The reference returns with the pivot slot still holding the
            // moved element (its data is then inconsistent); restore it.
            row[L] = pr;
            col[L] = pc;
            val[L] = pv;
            set_error("sort_coo exceeded %d levels", CLSPMV_MAX_LEVELS);
            return CLSPMV_E_SORT;
        }
        while (L < R) {
            while (L < R && greater(R, pr, pc)) --R;
            if (L < R) {
                row[L] = row[R];
                col[L] = col[R];
                val[L] = val[R];
                ++L;
            }
            while (L < R && less(L, pr, pc)) ++L;
            if (L < R) {
                row[R] = row[L];
                col[R] = col[L];
                val[R] = val[L];
                --R;
            }
        }
        row[L] = pr;
        col[L] = pc;
        val[L] = pv;
        lo[depth + 1] = L + 1;
        hi[depth + 1] = hi[depth];
        hi[depth] = L;
        ++depth;
    }
    return CLSPMV_OK;
}

// coo2csr (matrix_storage.h:1268-1307).
int clspmv_coo2csr(clspmv_coo_matrix* src, clspmv_csr_matrix* dst) {
    int rc = ensure_sorted(src);
    if (rc) return rc;
    std::vector<int> rp;
    if ((rc = row_ptr_of(src, rp))) return rc;
    const int nnz = src->matinfo.nnz;
    const int h = src->matinfo.height;
    dst->matinfo = src->matinfo;
    dst->csr_row_ptr = xmalloc<int>((size_t)h + 1);
    dst->csr_col_id = xmalloc<int>((size_t)nnz);
    dst->csr_data = xmalloc<float>((size_t)nnz);
    if (!dst->csr_row_ptr || !dst->csr_col_id || !dst->csr_data) {
        clspmv_free_csr(dst);
        return CLSPMV_E_NOMEM;
    }
    std::memcpy(dst->csr_row_ptr, rp.data(), sizeof(int) * ((size_t)h + 1));
    std::memcpy(dst->csr_col_id, src->coo_col_id, sizeof(int) * (size_t)nnz);
    std::memcpy(dst->csr_data, src->coo_data, sizeof(float) * (size_t)nnz);
    return CLSPMV_OK;
}

// coo2dia (matrix_storage.h:936-1014).  Diagonals are numbered in order of
// first appearance, then ordered by offset; data[d*L + row] = A[row][row+off].
// Refusal rule of the reference: when the k-th distinct diagonal (k = number
// already found) appears and k*L >= 2*nnz, it stops with dia_num = k.
// Rows >= L (only possible when height > width) overflow the reference's
// buffers; here they are rejected with CLSPMV_E_INVALID.
int clspmv_coo2dia(clspmv_coo_matrix* src, clspmv_dia_matrix* dst, int alignment) {
    dst->matinfo = src->matinfo;
    dst->dia_offsets = nullptr;
    dst->dia_data = nullptr;
    const int w = src->matinfo.width, h = src->matinfo.height, nnz = src->matinfo.nnz;
    dst->dia_length = std::min(w, h);
    const int L = clspmv_aligned_length(dst->dia_length, alignment);
    // Distinct offsets in order of first appearance.
    std::vector<int> first_order;
    {
        // offset in [-(h-1), w-1]; index by offset + h.
        std::vector<int> seen((size_t)w + (size_t)h + 1, -1);
        for (int k = 0; k < nnz; ++k) {
            int off = src->coo_col_id[k] - src->coo_row_id[k];
            int64_t idx = (int64_t)off + h;
            if (idx < 0 || idx >= (int64_t)seen.size()) {
                set_error("entry %d outside the matrix", k);
                return CLSPMV_E_INVALID;
            }
            if (seen[(size_t)idx] < 0) {
                int kdist = (int)first_order.size();
                if ((int64_t)kdist * L >= (int64_t)nnz * 2) {
                    dst->dia_num = kdist;
                    dst->dia_length_aligned = L;
                    set_error("diaCount %d length %d nnz %d: too much DIA padding", kdist, L, nnz);
                    return CLSPMV_E_PADDING;
                }
                seen[(size_t)idx] = kdist;
                first_order.push_back(off);
            }
        }
    }
    const int D = (int)first_order.size();
    std::vector<int> offs(first_order);
    std::sort(offs.begin(), offs.end());
    // slot of each offset after sorting (offsets are distinct).
    std::vector<int> slot_of((size_t)w + (size_t)h + 1, -1);
    for (int d = 0; d < D; ++d) slot_of[(size_t)(offs[d] + h)] = d;
    for (int k = 0; k < nnz; ++k)
        if (src->coo_row_id[k] >= L) {
            set_error("row %d >= diagonal length %d", src->coo_row_id[k], L);
            return CLSPMV_E_INVALID;
        }
    dst->dia_num = D;
    dst->dia_length_aligned = L;
    dst->dia_offsets = xmalloc<int>((size_t)D);
    dst->dia_data = xcalloc<float>((size_t)D * (size_t)L);
    if (!dst->dia_offsets || !dst->dia_data) {
        clspmv_free_dia(dst);
        return CLSPMV_E_NOMEM;
    }
    std::memcpy(dst->dia_offsets, offs.data(), sizeof(int) * (size_t)D);
    for (int k = 0; k < nnz; ++k) {
        int r = src->coo_row_id[k];
        int d = slot_of[(size_t)(src->coo_col_id[k] - r + h)];
        dst->dia_data[(size_t)d * L + r] = src->coo_data[k];
    }
    return CLSPMV_OK;
}

// coo2bdia (matrix_storage.h:1184-1236): DIA, then maximal runs of offsets
// that increase by exactly 1 become bands.  With no diagonal at all the
// reference reads offsets[0] of an empty array; here the result has 0 bands.
int clspmv_coo2bdia(clspmv_coo_matrix* src, clspmv_bdia_matrix* dst, int alignment) {
    dst->matinfo = src->matinfo;
    dst->bdia_offsets = nullptr;
    dst->bdia_bptr = nullptr;
    dst->bdia_data = nullptr;
    clspmv_dia_matrix dia;
    std::memset(&dia, 0, sizeof(dia));
    int rc = clspmv_coo2dia(src, &dia, alignment);
    if (rc) {
        clspmv_free_dia(&dia);
        return rc;
    }
    std::vector<int> boff, bcnt;
    for (int d = 0; d < dia.dia_num; ++d) {
        if (d > 0 && dia.dia_offsets[d] - dia.dia_offsets[d - 1] == 1) {
            bcnt.back()++;
        } else {
            boff.push_back(dia.dia_offsets[d]);
            bcnt.push_back(1);
        }
    }
    const int nb = (int)boff.size();
    dst->bdia_band_num = nb;
    dst->bdia_length = dia.dia_length;
    dst->bdia_length_aligned = dia.dia_length_aligned;
    dst->bdia_offsets = xmalloc<int>((size_t)nb);
    dst->bdia_bptr = xmalloc<int>((size_t)nb + 1);
    dst->bdia_data = dia.dia_data;  // same array as DIA (memcpy in the reference)
    dia.dia_data = nullptr;
    clspmv_free_dia(&dia);
    if (!dst->bdia_offsets || !dst->bdia_bptr) {
        clspmv_free_bdia(dst);
        return CLSPMV_E_NOMEM;
    }
    dst->bdia_bptr[0] = 0;
    for (int b = 0; b < nb; ++b) {
        dst->bdia_offsets[b] = boff[b];
        dst->bdia_bptr[b + 1] = dst->bdia_bptr[b] + bcnt[b];
    }
    return CLSPMV_OK;
}

// coo2ell (matrix_storage.h:1340-1394): slot-major arrays of height
// H = align(height, alignment); padding slots of a row repeat the row's last
// column (0 for an empty row) with value 0; rows >= height hold col 0, value 0.
int clspmv_coo2ell(clspmv_coo_matrix* src, clspmv_ell_matrix* dst, int alignment, int ellnum) {
    int rc = ensure_sorted(src);
    if (rc) return rc;
    std::vector<int> rp;
    if ((rc = row_ptr_of(src, rp))) return rc;
    const int h = src->matinfo.height;
    dst->matinfo = src->matinfo;
    if (ellnum == 0)
        for (int r = 0; r < h; ++r) ellnum = std::max(ellnum, rp[(size_t)r + 1] - rp[r]);
    for (int r = 0; r < h; ++r)
        if (rp[(size_t)r + 1] - rp[r] > ellnum) {
            set_error("row %d has %d > ell_num %d entries", r, rp[(size_t)r + 1] - rp[r], ellnum);
            return CLSPMV_E_INVALID;
        }
    const int H = clspmv_aligned_length(h, alignment);
    dst->ell_num = ellnum;
    dst->ell_height_aligned = H;
    size_t total = (size_t)H * (size_t)ellnum;
    dst->ell_col_id = xcalloc<int>(total);
    dst->ell_data = xcalloc<float>(total);
    if (!dst->ell_col_id || !dst->ell_data) {
        clspmv_free_ell(dst);
        return CLSPMV_E_NOMEM;
    }
#pragma omp parallel for schedule(static) num_threads(omp_threads())
    for (int r = 0; r < h; ++r) {
        int s = rp[r], e = rp[(size_t)r + 1];
        int last = 0;
        for (int k = s; k < e; ++k) {
            size_t idx = (size_t)(k - s) * H + r;
            dst->ell_col_id[idx] = src->coo_col_id[k];
            dst->ell_data[idx] = src->coo_data[k];
            last = src->coo_col_id[k];
        }
        for (int j = e - s; j < ellnum; ++j) dst->ell_col_id[(size_t)j * H + r] = last;
    }
    return CLSPMV_OK;
}

// coo2sell (matrix_storage.h:1397-1513).  Slice s holds rows
// [s*S, s*S+S); its width is the longest row in it; element (row, slot) sits
// at slice_ptr[s] + slot*S + row%S.  Padding slot k of a row copies the
// column of slot k-1 of the same row, and slot 0 of an empty row copies the
// slice's element 0 (first column of the slice's first row, or 0); rows past
// the matrix height in the last slice follow the same rule.  A slice of width
// 0 stores nothing (the reference writes into malloc(0) there).
int clspmv_coo2sell(clspmv_coo_matrix* src, clspmv_sell_matrix* dst, unsigned int slice_height) {
    if (slice_height == 0) return CLSPMV_E_INVALID;
    // The reference assumes a sorted source here; an unsorted one is sorted.
    int rc = ensure_sorted(src);
    if (rc) return rc;
    std::vector<int> rp;
    if ((rc = row_ptr_of(src, rp))) return rc;
    const int h = src->matinfo.height;
    const int S = (int)slice_height;
    const int ns = (h + S - 1) / S;
    dst->matinfo = src->matinfo;
    dst->sell_slice_height = slice_height;
    dst->sell_slice_num = ns;
    dst->sell_slice_ptr = xmalloc<int>((size_t)ns + 1);
    if (!dst->sell_slice_ptr) return CLSPMV_E_NOMEM;
    dst->sell_slice_ptr[0] = 0;
    std::vector<int> width((size_t)ns, 0);
    for (int s = 0; s < ns; ++s) {
        int wmax = 0;
        for (int r = s * S; r < std::min(h, s * S + S); ++r) wmax = std::max(wmax, rp[(size_t)r + 1] - rp[r]);
        width[s] = wmax;
        int64_t next = (int64_t)dst->sell_slice_ptr[s] + (int64_t)wmax * S;
        if (next > INT32_MAX) {
            clspmv_free_sell(dst);
            return CLSPMV_E_TOO_LARGE;
        }
        dst->sell_slice_ptr[s + 1] = (int)next;
    }
    size_t total = (size_t)dst->sell_slice_ptr[ns];
    dst->sell_col_id = xcalloc<int>(total);
    dst->sell_data = xcalloc<float>(total);
    if (!dst->sell_col_id || !dst->sell_data) {
        clspmv_free_sell(dst);
        return CLSPMV_E_NOMEM;
    }
#pragma omp parallel for schedule(static) num_threads(omp_threads())
    for (int s = 0; s < ns; ++s) {
        const int wmax = width[s];
        if (wmax == 0) continue;
        int* col = dst->sell_col_id + dst->sell_slice_ptr[s];
        float* dat = dst->sell_data + dst->sell_slice_ptr[s];
        const int off = s * S;
        col[0] = 0;
        // Row order matters: padding reads entries written by earlier rows.
        for (int r = off; r < off + S; ++r) {
            int len = 0;
            if (r < h) {
                int st = rp[r];
                len = rp[(size_t)r + 1] - st;
                for (int k = 0; k < len; ++k) {
                    col[k * S + (r - off)] = src->coo_col_id[st + k];
                    dat[k * S + (r - off)] = src->coo_data[st + k];
                }
            }
            for (int k = len; k < wmax; ++k) {
                int id = k * S + (r - off);
                col[id] = (id >= S) ? col[id - S] : col[0];
                dat[id] = 0.0f;
            }
        }
    }
    return CLSPMV_OK;
}

// coo2b4csr (matrix_storage.h:1740-1897): blocks of bh x bw (bw in {4,8}),
// block column stored in float4 units (bcol/4); with N = align(4*nblocks,
// alignment) the value of block i at (h, w4*4+w) lives at
//   data[w4*(N*bh) + h*N + 4*i + w].
int clspmv_coo2b4csr(clspmv_coo_matrix* src, clspmv_b4csr_matrix* dst, unsigned int bwidth,
                     unsigned int bheight, int alignment) {
    if (bwidth % 4 != 0 || bwidth == 0 || bheight == 0) return CLSPMV_E_INVALID;
    int rc = ensure_sorted(src);
    if (rc) return rc;
    std::vector<int> rp;
    if ((rc = row_ptr_of(src, rp))) return rc;
    BlockRows B;
    extract_blocks(src, rp, bwidth, bheight, B);
    const size_t nb = B.bcol.size();
    const unsigned bs = bwidth * bheight;
    if (nb > (size_t)(CLSPMV_MAX_MEM_OBJ / (bs * sizeof(float)))) {
        set_error("BCSR too large blocknum %zu bwidth %u bheight %u", nb, bwidth, bheight);
        return CLSPMV_E_TOO_LARGE;
    }
    const int brows = (int)B.rowptr.size() - 1;
    dst->matinfo = src->matinfo;
    dst->b4csr_bwidth = bwidth;
    dst->b4csr_bheight = bheight;
    dst->b4csr_row_num = brows;
    const int N = clspmv_aligned_length((int)nb * 4, alignment);
    const unsigned w4n = bwidth / 4;
    dst->b4csr_aligned_size = N;
    dst->b4csr_block_num = (int)nb;
    dst->b4csr_row_ptr = xmalloc<int>((size_t)brows + 1);
    dst->b4csr_col_id = xmalloc<int>(nb);
    dst->b4csr_data = xcalloc<float>((size_t)N * bheight * w4n);
    if (!dst->b4csr_row_ptr || !dst->b4csr_col_id || !dst->b4csr_data) {
        clspmv_free_b4csr(dst);
        return CLSPMV_E_NOMEM;
    }
    std::memcpy(dst->b4csr_row_ptr, B.rowptr.data(), sizeof(int) * ((size_t)brows + 1));
    const size_t plane = (size_t)N * bheight;
#pragma omp parallel for schedule(static) num_threads(omp_threads())
    for (int64_t i = 0; i < (int64_t)nb; ++i) {
        dst->b4csr_col_id[i] = B.bcol[i] / 4;
        const float* blk = &B.data[(size_t)i * bs];
        for (unsigned hh = 0; hh < bheight; ++hh)
            for (unsigned w4 = 0; w4 < w4n; ++w4)
                for (unsigned w = 0; w < 4; ++w)
                    dst->b4csr_data[w4 * plane + (size_t)hh * N + (size_t)i * 4 + w] = blk[hh * bwidth + w4 * 4 + w];
    }
    return CLSPMV_OK;
}

// coo2b4ell (matrix_storage.h:2009-2200): slot-major blocked ELL.
// Hc = align(brows, alignment) column slots per ell slot, F = align(4*brows,
// alignment); value (slot j, block row r, h, w4*4+w) lives at
//   data[j*(F*w4n*bh) + h*F + w4*(F*bh) + 4*r + w].
// Padding slots repeat the block row's last block column with zero values.
// Entries of block rows >= brows inside the aligned region are left
// uninitialised by the reference; here they are 0.
int clspmv_coo2b4ell(clspmv_coo_matrix* src, clspmv_b4ell_matrix* dst, unsigned int bwidth,
                     unsigned int bheight, int alignment, int ellnum) {
    if (bwidth % 4 != 0 || bwidth == 0 || bheight == 0) return CLSPMV_E_INVALID;
    int rc = ensure_sorted(src);
    if (rc) return rc;
    std::vector<int> rp;
    if ((rc = row_ptr_of(src, rp))) return rc;
    BlockRows B;
    extract_blocks(src, rp, bwidth, bheight, B);
    const int brows = (int)B.rowptr.size() - 1;
    const unsigned bs = bwidth * bheight;
    if (ellnum == 0)
        for (int r = 0; r < brows; ++r) ellnum = std::max(ellnum, B.rowptr[(size_t)r + 1] - B.rowptr[r]);
    if ((size_t)brows * (size_t)ellnum > (size_t)(CLSPMV_MAX_MEM_OBJ / (bs * sizeof(float)))) {
        set_error("BELL too large ellnum %d bwidth %u bheight %u", ellnum, bwidth, bheight);
        return CLSPMV_E_TOO_LARGE;
    }
    for (int r = 0; r < brows; ++r)
        if (B.rowptr[(size_t)r + 1] - B.rowptr[r] > ellnum) return CLSPMV_E_INVALID;
    dst->matinfo = src->matinfo;
    dst->b4ell_bwidth = bwidth;
    dst->b4ell_bheight = bheight;
    dst->b4ell_row_num = brows;
    dst->b4ell_block_num = ellnum;
    const int Hc = clspmv_aligned_length(brows, alignment);
    const int F = clspmv_aligned_length(4 * brows, alignment);
    const unsigned w4n = bwidth / 4;
    dst->b4ell_height_aligned = Hc;
    dst->b4ell_float4_aligned = F;
    dst->b4ell_col_id = xcalloc<int>((size_t)Hc * ellnum);
    dst->b4ell_data = xcalloc<float>((size_t)F * bheight * w4n * ellnum);
    if (!dst->b4ell_col_id || !dst->b4ell_data) {
        clspmv_free_b4ell(dst);
        return CLSPMV_E_NOMEM;
    }
    const size_t slot_stride = (size_t)F * w4n * bheight;
    const size_t w4_stride = (size_t)F * bheight;
#pragma omp parallel for schedule(static) num_threads(omp_threads())
    for (int r = 0; r < brows; ++r) {
        int s = B.rowptr[r], e = B.rowptr[(size_t)r + 1];
        int last = 0;
        for (int j = s; j < e; ++j) {
            int c4 = B.bcol[j] / 4;
            dst->b4ell_col_id[r + (size_t)(j - s) * Hc] = c4;
            last = c4;
            const float* blk = &B.data[(size_t)j * bs];
            for (unsigned hh = 0; hh < bheight; ++hh)
                for (unsigned w4 = 0; w4 < w4n; ++w4)
                    for (unsigned w = 0; w < 4; ++w)
                        dst->b4ell_data[(size_t)(j - s) * slot_stride + (size_t)hh * F + w4 * w4_stride + (size_t)r * 4 + w] =
                            blk[hh * bwidth + w4 * 4 + w];
        }
        for (int j = e - s; j < ellnum; ++j) dst->b4ell_col_id[r + (size_t)j * Hc] = last;
    }
    return CLSPMV_OK;
}

// coo2sbell (matrix_storage.h:2202-2409): slices of S block rows, per-slice
// width; col of (slice s, row-in-slice t, slot j) at slice_ptr[s] + t + j*S,
// value (h, w4*4+w) at slice_ptr[s]*bw*bh + j*(S*4*w4n*bh) + h*(S*4)
//                      + w4*(S*4*bh) + 4*t + w.
// On a MAX_MEM_OBJ refusal the reference frees slice_ptr but leaves the
// pointer set; here it is reset to NULL.
int clspmv_coo2sbell(clspmv_coo_matrix* src, clspmv_sbell_matrix* dst, unsigned int bwidth,
                     unsigned int bheight, unsigned int slice_height) {
    if (bwidth % 4 != 0 || bwidth == 0 || bheight == 0 || slice_height == 0) return CLSPMV_E_INVALID;
    int rc = ensure_sorted(src);
    if (rc) return rc;
    std::vector<int> rp;
    if ((rc = row_ptr_of(src, rp))) return rc;
    BlockRows B;
    extract_blocks(src, rp, bwidth, bheight, B);
    const int brows = (int)B.rowptr.size() - 1;
    const unsigned bs = bwidth * bheight;
    const int S = (int)slice_height;
    const int ns = (brows + S - 1) / S;
    dst->matinfo = src->matinfo;
    dst->sbell_bwidth = bwidth;
    dst->sbell_bheight = bheight;
    dst->sbell_slice_num = ns;
    dst->sbell_row_num = brows;
    dst->sbell_slice_height = slice_height;
    dst->sbell_col_id = nullptr;
    dst->sbell_data = nullptr;
    dst->sbell_slice_ptr = xmalloc<int>((size_t)ns + 1);
    if (!dst->sbell_slice_ptr) return CLSPMV_E_NOMEM;
    std::vector<int> width((size_t)ns, 0);
    dst->sbell_slice_ptr[0] = 0;
    int64_t acc = 0;
    for (int s = 0; s < ns; ++s) {
        int wmax = 0;
        for (int r = s * S; r < std::min(brows, s * S + S); ++r) wmax = std::max(wmax, B.rowptr[(size_t)r + 1] - B.rowptr[r]);
        width[s] = wmax;
        acc += (int64_t)wmax * S;
        dst->sbell_slice_ptr[s + 1] = (int)std::min<int64_t>(acc, INT32_MAX);
    }
    if (acc > (int64_t)(CLSPMV_MAX_MEM_OBJ / (bs * sizeof(float)))) {
        std::free(dst->sbell_slice_ptr);
        dst->sbell_slice_ptr = nullptr;
        set_error("SBELL too large totalsize %lld bwidth %u bheight %u", (long long)acc, bwidth, bheight);
        return CLSPMV_E_TOO_LARGE;
    }
    const size_t total = (size_t)acc;
    dst->sbell_col_id = xcalloc<int>(total);
    dst->sbell_data = xcalloc<float>(total * bs);
    if (!dst->sbell_col_id || !dst->sbell_data) {
        clspmv_free_sbell(dst);
        return CLSPMV_E_NOMEM;
    }
    const unsigned w4n = bwidth / 4;
    const size_t slot_stride = (size_t)S * w4n * bheight * 4;
    const size_t w4_stride = (size_t)S * bheight * 4;
    const size_t h_stride = (size_t)S * 4;
#pragma omp parallel for schedule(static) num_threads(omp_threads())
    for (int r = 0; r < brows; ++r) {
        int s = B.rowptr[r], e = B.rowptr[(size_t)r + 1];
        int sl = r / S, t = r % S;
        size_t colstart = (size_t)dst->sbell_slice_ptr[sl];
        size_t datastart = colstart * bs;
        int last = 0;
        for (int j = s; j < e; ++j) {
            int c4 = B.bcol[j] / 4;
            dst->sbell_col_id[colstart + t + (size_t)(j - s) * S] = c4;
            last = c4;
            const float* blk = &B.data[(size_t)j * bs];
            for (unsigned hh = 0; hh < bheight; ++hh)
                for (unsigned w4 = 0; w4 < w4n; ++w4)
                    for (unsigned w = 0; w < 4; ++w)
                        dst->sbell_data[datastart + (size_t)(j - s) * slot_stride + hh * h_stride + w4 * w4_stride +
                                        (size_t)t * 4 + w] = blk[hh * bwidth + w4 * 4 + w];
        }
        for (int j = e - s; j < width[sl]; ++j) dst->sbell_col_id[colstart + t + (size_t)j * S] = last;
    }
    return CLSPMV_OK;
}

// Native r x c BCSR (no reference counterpart).  Block extraction is the same
// as the reference's blocked formats (block column = col - col % bwidth).
int clspmv_coo2bsr(clspmv_coo_matrix* src, clspmv_bsr_matrix* dst, unsigned int bwidth, unsigned int bheight) {
    if (bwidth == 0 || bheight == 0 || bwidth > 8 || bheight > 8) return CLSPMV_E_INVALID;
    int rc = ensure_sorted(src);
    if (rc) return rc;
    std::vector<int> rp;
    if ((rc = row_ptr_of(src, rp))) return rc;
    BlockRows B;
    extract_blocks(src, rp, bwidth, bheight, B);
    const int brows = (int)B.rowptr.size() - 1;
    const size_t nb = B.bcol.size();
    const unsigned bs = bwidth * bheight;
    const int stride = clspmv_aligned_length((int)nb, 64);
    dst->matinfo = src->matinfo;
    dst->bsr_bwidth = bwidth;
    dst->bsr_bheight = bheight;
    dst->bsr_row_num = brows;
    dst->bsr_block_num = (int)nb;
    dst->bsr_block_stride = stride;
    dst->bsr_row_ptr = xmalloc<int>((size_t)brows + 1);
    dst->bsr_col_id = xmalloc<int>(nb);
    dst->bsr_data = xcalloc<float>((size_t)stride * bs);
    if (!dst->bsr_row_ptr || !dst->bsr_col_id || !dst->bsr_data) {
        clspmv_free_bsr(dst);
        return CLSPMV_E_NOMEM;
    }
    std::memcpy(dst->bsr_row_ptr, B.rowptr.data(), sizeof(int) * ((size_t)brows + 1));
    std::memcpy(dst->bsr_col_id, B.bcol.data(), sizeof(int) * nb);
#pragma omp parallel for schedule(static) num_threads(omp_threads())
    for (int64_t i = 0; i < (int64_t)nb; ++i)
        for (unsigned e = 0; e < bs; ++e) dst->bsr_data[(size_t)e * stride + (size_t)i] = B.data[(size_t)i * bs + e];
    return CLSPMV_OK;
}

// pad_csr (A/src/util.cpp:31-72): every row is padded to a multiple of
// `alignment` entries, repeating its last column with value 0.
int clspmv_pad_csr(const clspmv_csr_matrix* src, clspmv_csr_matrix* dst, int alignment) {
    const int h = src->matinfo.height;
    dst->matinfo.height = h;
    dst->matinfo.width = src->matinfo.width;
    dst->csr_row_ptr = xmalloc<int>((size_t)h + 1);
    if (!dst->csr_row_ptr) return CLSPMV_E_NOMEM;
    dst->csr_row_ptr[0] = 0;
    for (int r = 0; r < h; ++r) {
        int len = src->csr_row_ptr[r + 1] - src->csr_row_ptr[r];
        dst->csr_row_ptr[r + 1] = dst->csr_row_ptr[r] + clspmv_aligned_length(len, alignment);
    }
    const int total = dst->csr_row_ptr[h];
    dst->matinfo.nnz = total;
    dst->csr_col_id = xmalloc<int>((size_t)total);
    dst->csr_data = xmalloc<float>((size_t)total);
    if (!dst->csr_col_id || !dst->csr_data) {
        clspmv_free_csr(dst);
        return CLSPMV_E_NOMEM;
    }
    for (int r = 0; r < h; ++r) {
        int s = src->csr_row_ptr[r], e = src->csr_row_ptr[r + 1];
        int o = dst->csr_row_ptr[r], oe = dst->csr_row_ptr[r + 1];
        for (int k = s; k < e; ++k) {
            dst->csr_col_id[o + (k - s)] = src->csr_col_id[k];
            dst->csr_data[o + (k - s)] = src->csr_data[k];
        }
        int last = (e > s) ? src->csr_col_id[e - 1] : 0;
        for (int k = o + (e - s); k < oe; ++k) {
            dst->csr_col_id[k] = last;
            dst->csr_data[k] = 0.0f;
        }
    }
    return CLSPMV_OK;
}

}  // extern "C"
