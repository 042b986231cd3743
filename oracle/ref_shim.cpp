// ref_shim.cpp -- extern "C" entry points into the reference's own CPU code.
//
// TEST INFRASTRUCTURE ONLY (see spmv_oracle.c).  Built by oracle/Makefile
// into oracle/_ref/libclspmv_ref.so together with the reference's
// unmodified sources A/src/util.cpp, A/src/fileio.cpp, A/src/analyze.cpp and
// headers A/include/*.h, read in place from the reference tree.  It lets the
// tests compare libclspmv's conversions, reader and analyzer with the
// reference's compiled implementation on the same inputs.
//
// The reference's structs (matrix_storage.h, <int,int,float>) have the same
// layout as the C structs of include/clspmv.h; that is asserted below so the
// Python side can pass the same ctypes structs to both libraries.
#include <cstddef>
#include <cstdlib>
#include <cstring>

#include "analyze.h"
#include "fileio.h"
#include "matrix_storage.h"
#include "spmv_serial.h"
#include "util.h"

#include "clspmv.h"

static_assert(sizeof(coo_matrix<int, float>) == sizeof(clspmv_coo_matrix), "coo layout");
static_assert(sizeof(csr_matrix<int, float>) == sizeof(clspmv_csr_matrix), "csr layout");
static_assert(sizeof(dia_matrix<int, int, float>) == sizeof(clspmv_dia_matrix), "dia layout");
static_assert(sizeof(bdia_matrix<int, int, float>) == sizeof(clspmv_bdia_matrix), "bdia layout");
static_assert(sizeof(ell_matrix<int, float>) == sizeof(clspmv_ell_matrix), "ell layout");
static_assert(sizeof(sell_matrix<int, float>) == sizeof(clspmv_sell_matrix), "sell layout");
static_assert(sizeof(b4csr_matrix<int, float>) == sizeof(clspmv_b4csr_matrix), "b4csr layout");
static_assert(sizeof(b4ell_matrix<int, float>) == sizeof(clspmv_b4ell_matrix), "b4ell layout");
static_assert(sizeof(sbell_matrix<int, float>) == sizeof(clspmv_sbell_matrix), "sbell layout");
typedef b4ell_matrix<int, float> RefB4Ell;
typedef sbell_matrix<int, float> RefSbell;
typedef sell_matrix<int, float> RefSell;
static_assert(offsetof(RefB4Ell, b4ell_data) == offsetof(clspmv_b4ell_matrix, b4ell_data), "b4ell");
static_assert(offsetof(RefSbell, sbell_data) == offsetof(clspmv_sbell_matrix, sbell_data), "sbell");
static_assert(offsetof(RefSell, sell_data) == offsetof(clspmv_sell_matrix, sell_data), "sell");

typedef coo_matrix<int, float> Coo;

extern "C" {

int ref_sort_coo(void* coo) { return sort_coo<int, float>((Coo*)coo) ? 1 : 0; }

int ref_coo2csr(void* coo, void* out) {
    coo2csr<int, float>((Coo*)coo, (csr_matrix<int, float>*)out);
    return 1;
}

int ref_coo2dia(void* coo, void* out, int alignment) {
    return coo2dia<int, int, float>((Coo*)coo, (dia_matrix<int, int, float>*)out, alignment) ? 1 : 0;
}

int ref_coo2bdia(void* coo, void* out, int alignment) {
    return coo2bdia<int, int, float>((Coo*)coo, (bdia_matrix<int, int, float>*)out, alignment) ? 1 : 0;
}

int ref_coo2ell(void* coo, void* out, int alignment, int ellnum) {
    coo2ell<int, float>((Coo*)coo, (ell_matrix<int, float>*)out, alignment, ellnum);
    return 1;
}

int ref_coo2sell(void* coo, void* out, unsigned slice_height) {
    coo2sell<int, float>((Coo*)coo, (sell_matrix<int, float>*)out, slice_height);
    return 1;
}

int ref_coo2b4csr(void* coo, void* out, unsigned bw, unsigned bh, int alignment) {
    return coo2b4csr<int, float>((Coo*)coo, (b4csr_matrix<int, float>*)out, bw, bh, alignment) ? 1 : 0;
}

int ref_coo2b4ell(void* coo, void* out, unsigned bw, unsigned bh, int alignment, int ellnum) {
    return coo2b4ell<int, float>((Coo*)coo, (b4ell_matrix<int, float>*)out, bw, bh, alignment, ellnum) ? 1 : 0;
}

int ref_coo2sbell(void* coo, void* out, unsigned bw, unsigned bh, unsigned slice_height) {
    return coo2sbell<int, float>((Coo*)coo, (sbell_matrix<int, float>*)out, bw, bh, slice_height) ? 1 : 0;
}

void ref_pad_csr(void* csr, void* out, int alignment) {
    pad_csr((csr_matrix<int, float>*)csr, (csr_matrix<int, float>*)out, alignment);
}

void ref_read_mmf(const char* filename, void* coo) { ReadMMF((char*)filename, (Coo*)coo); }

void ref_spmv_only(void* coo, float* vec, float* res) { spmv_only((Coo*)coo, vec, res); }

// The reference's free_* only free the arrays (malloc'd inside the
// conversions), exactly like libc free.
void ref_free(void* p) { std::free(p); }

}  // extern "C"

// Functions of A/src/analyze.cpp with external linkage that analyze.h does
// not declare.
void update_info(coo_matrix<int, float>& coomat, maxflop_info& info, benchmark_all& bench);
double find_overhead(std::vector<overhead_entry>& benchover, int groupnum, int threadnum);
void ReadBench(benchmark_all& bench);

#include "clspmv_cocktail.h"
static_assert(sizeof(cocktail<int, int, float>) == sizeof(clspmv_cocktail), "cocktail layout");
static_assert(sizeof(maxflop_info) == sizeof(clspmv_maxflop_info), "maxflop_info layout");

static benchmark_all* g_bench = nullptr;

extern "C" {

// analyze_matrix with the tables of $CLSPMVPATH/benchmark.
void ref_analyze_matrix(void* coo, void* mat, float dia_mult, int partial, int full) {
    analyze_matrix(*(Coo*)coo, *(cocktail<int, int, float>*)mat, dia_mult, partial != 0, full != 0);
}

// update_info for a (rows, nnz) point: every maxflop_* of the reference.
void ref_maxflop(int rows, int nnz, void* out) {
    if (!g_bench) {
        g_bench = new benchmark_all();
        ReadBench(*g_bench);
    }
    Coo m;
    init_coo_matrix(m);
    m.matinfo.height = rows;
    m.matinfo.width = rows;
    m.matinfo.nnz = nnz;
    maxflop_info info;
    std::memset(&info, 0, sizeof(info));
    update_info(m, info, *g_bench);
    std::memcpy(out, &info, sizeof(info));
}

double ref_find_overhead(int groupnum, int threadnum) {
    if (!g_bench) {
        g_bench = new benchmark_all();
        ReadBench(*g_bench);
    }
    return find_overhead(g_bench->overhead, groupnum, threadnum);
}

}  // extern "C"
