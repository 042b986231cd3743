// Internal helpers shared by the host translation units of libclspmv.
#pragma once
#include <cstdarg>
#include <cstdio>
#include <vector>

#include "clspmv.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace clspmv {

void set_error(const char* fmt, ...);

inline int omp_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
inline int omp_tid() {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

int row_ptr_of(const clspmv_coo_matrix* m, std::vector<int>& rp);
int check_cols(const clspmv_coo_matrix* m);
int ensure_sorted(clspmv_coo_matrix* m);

}  // namespace clspmv
