// Matrix Market and binary COO I/O.
//
// clspmv_read_mmf follows ReadMMF (A/src/fileio.cpp:10-109) token for token,
// including its quirks:
//   * the header is read as whitespace-separated tokens; only the 4th token is
//     tested ("real" -> values are read, anything else -> every value is 1.0
//     and no value token is consumed, so an "integer" file is mis-parsed
//     exactly as the reference mis-parses it);
//   * the 5th token "symmetric" mirrors every off-diagonal entry (skew
//     matrices are not negated); duplicates are kept;
//   * comment lines are skipped token-wise ('%' token -> rest of line);
//   * the result is sorted with sort_coo semantics.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "clspmv.h"
#include "internal.hpp"

using namespace clspmv;

namespace {

struct TokenReader {
    FILE* f;
    // Reads one whitespace-delimited token like fscanf("%s").
    bool token(std::string& out) {
        out.clear();
        int c;
        do {
            c = std::fgetc(f);
            if (c == EOF) return false;
        } while (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f');
        while (c != EOF && !(c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f')) {
            out.push_back((char)c);
            c = std::fgetc(f);
        }
        if (c != EOF) std::ungetc(c, f);
        return true;
    }
    // fgets(buf, 1025, f) equivalent: consume up to newline (at most 1024 chars).
    void rest_of_line() {
        int n = 0, c;
        while (n < 1024 && (c = std::fgetc(f)) != EOF) {
            ++n;
            if (c == '\n') break;
        }
    }
};

}  // namespace

extern "C" {

int clspmv_read_mmf(const char* filename, clspmv_coo_matrix* m) {
    FILE* f = std::fopen(filename, "r");
    if (!f) {
        set_error("cannot open %s", filename);
        return CLSPMV_E_IO;
    }
    TokenReader tr{f};
    std::string tok;
    bool ok = true;
    for (int i = 0; i < 4 && ok; ++i) ok = tr.token(tok);
    bool is_real = ok && tok == "real";
    ok = ok && tr.token(tok);
    bool is_sym = ok && tok == "symmetric";
    int height = 0, width = 0;
    long long nnz = 0;
    while (ok) {
        ok = tr.token(tok);
        if (!ok) break;
        if (tok[0] != '%') {
            height = std::atoi(tok.c_str());
            break;
        }
        tr.rest_of_line();
    }
    if (!ok || std::fscanf(f, "%d %lld", &width, &nnz) != 2 || nnz < 0 || nnz > INT32_MAX) {
        std::fclose(f);
        set_error("malformed Matrix Market header in %s", filename);
        return CLSPMV_E_IO;
    }
    std::vector<int> rows((size_t)nnz), cols((size_t)nnz);
    std::vector<float> vals((size_t)nnz, 1.0f);
    long long diag = 0;
    for (long long i = 0; i < nnz; ++i) {
        int r = 0, c = 0;
        if (std::fscanf(f, "%d %d", &r, &c) != 2) {
            std::fclose(f);
            set_error("truncated entry list in %s at entry %lld", filename, i);
            return CLSPMV_E_IO;
        }
        rows[(size_t)i] = r - 1;
        cols[(size_t)i] = c - 1;
        if (is_real) {
            double v = 0.0;
            if (std::fscanf(f, "%lf", &v) != 1) {
                std::fclose(f);
                set_error("missing value in %s at entry %lld", filename, i);
                return CLSPMV_E_IO;
            }
            vals[(size_t)i] = (float)v;
        }
        if (rows[(size_t)i] == cols[(size_t)i]) ++diag;
    }
    std::fclose(f);
    long long total = is_sym ? 2 * nnz - diag : nnz;
    if (total > INT32_MAX) return CLSPMV_E_TOO_LARGE;
    int rc = clspmv_alloc_coo(m, width, height, (int)total);
    if (rc) return rc;
    size_t k = 0;
    for (long long i = 0; i < nnz; ++i) {
        m->coo_row_id[k] = rows[(size_t)i];
        m->coo_col_id[k] = cols[(size_t)i];
        m->coo_data[k] = vals[(size_t)i];
        ++k;
        if (is_sym && rows[(size_t)i] != cols[(size_t)i]) {
            m->coo_row_id[k] = cols[(size_t)i];
            m->coo_col_id[k] = rows[(size_t)i];
            m->coo_data[k] = vals[(size_t)i];
            ++k;
        }
    }
    // The reference sorts a symmetric expansion unconditionally and a general
    // matrix only when unsorted (the quicksort may reorder equal keys even in
    // a sorted array, so the distinction matters for duplicate entries).
    if (is_sym || !clspmv_if_sorted_coo(m)) {
        rc = clspmv_sort_coo(m);
        if (rc) return rc;
    }
    return CLSPMV_OK;
}

int clspmv_write_mmf(const char* filename, const clspmv_coo_matrix* m) {
    FILE* f = std::fopen(filename, "w");
    if (!f) return CLSPMV_E_IO;
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(f, "%d %d %d\n", m->matinfo.height, m->matinfo.width, m->matinfo.nnz);
    for (int k = 0; k < m->matinfo.nnz; ++k)
        std::fprintf(f, "%d %d %.9g\n", m->coo_row_id[k] + 1, m->coo_col_id[k] + 1, (double)m->coo_data[k]);
    return std::fclose(f) == 0 ? CLSPMV_OK : CLSPMV_E_IO;
}

int clspmv_write_coo_bin(const char* filename, const clspmv_coo_matrix* m) {
    FILE* f = std::fopen(filename, "wb");
    if (!f) return CLSPMV_E_IO;
    size_t n = (size_t)m->matinfo.nnz;
    bool ok = std::fwrite(&m->matinfo.width, sizeof(int), 1, f) == 1 &&
              std::fwrite(&m->matinfo.height, sizeof(int), 1, f) == 1 &&
              std::fwrite(&m->matinfo.nnz, sizeof(int), 1, f) == 1 &&
              std::fwrite(m->coo_row_id, sizeof(int), n, f) == n && std::fwrite(m->coo_col_id, sizeof(int), n, f) == n &&
              std::fwrite(m->coo_data, sizeof(float), n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? CLSPMV_OK : CLSPMV_E_IO;
}

int clspmv_read_coo_bin(const char* filename, clspmv_coo_matrix* m) {
    FILE* f = std::fopen(filename, "rb");
    if (!f) return CLSPMV_E_IO;
    int hdr[3];
    if (std::fread(hdr, sizeof(int), 3, f) != 3 || hdr[0] < 0 || hdr[1] < 0 || hdr[2] < 0) {
        std::fclose(f);
        return CLSPMV_E_IO;
    }
    int rc = clspmv_alloc_coo(m, hdr[0], hdr[1], hdr[2]);
    if (rc) {
        std::fclose(f);
        return rc;
    }
    size_t n = (size_t)hdr[2];
    bool ok = std::fread(m->coo_row_id, sizeof(int), n, f) == n && std::fread(m->coo_col_id, sizeof(int), n, f) == n &&
              std::fread(m->coo_data, sizeof(float), n, f) == n;
    std::fclose(f);
    if (!ok) {
        clspmv_free_coo(m);
        return CLSPMV_E_IO;
    }
    return CLSPMV_OK;
}

}  // extern "C"
