// Seeded synthetic matrix generators for the benchmark and test matrices.
//
// PRNG: SplitMix64 seeds one xoshiro256** stream per row (stream key =
// seed, row), so generation is deterministic and independent of the thread
// count.  Uniform floats use the top 24 bits; normals use Box-Muller in
// double precision.  Every generator emits a row-major sorted COO with
// distinct (row, col) pairs (the reference's quicksort is not stable, so
// duplicates would make conversions order-dependent).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "clspmv.h"
#include "internal.hpp"

using namespace clspmv;

namespace {

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Xoshiro {
    uint64_t s[4];
    Xoshiro(uint64_t seed, uint64_t stream) {
        uint64_t sm = seed ^ (stream * 0xD1B54A32D192ED03ull) ^ 0x5851F42D4C957F2Dull;
        for (int i = 0; i < 4; ++i) s[i] = splitmix64(sm);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform double in [0,1) with 53 bits
    inline double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform integer in [0, n)
    inline uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
    inline float normal() {
        double u1 = 1.0 - uniform();  // (0,1]
        double u2 = uniform();
        return (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
};

// Draws `k` distinct integers in [lo, hi) in ascending order.
void distinct_sorted(Xoshiro& g, int lo, int hi, int k, std::vector<int>& out) {
    out.clear();
    const int span = hi - lo;
    if (k <= 0 || span <= 0) return;
    if (k >= span) {
        for (int c = lo; c < hi; ++c) out.push_back(c);
        return;
    }
    while ((int)out.size() < k) {
        int need = k - (int)out.size();
        for (int i = 0; i < need; ++i) out.push_back(lo + (int)g.below((uint32_t)span));
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    }
}

// Fills a COO from per-row lengths and a per-row generator callback.
template <class RowFn>
int build_rows(clspmv_coo_matrix* m, int width, int height, const std::vector<int>& len, RowFn fn) {
    std::vector<int64_t> start((size_t)height + 1, 0);
    for (int r = 0; r < height; ++r) start[(size_t)r + 1] = start[r] + len[r];
    if (start[height] > INT32_MAX) {
        set_error("generated matrix has %lld > 2^31-1 entries", (long long)start[height]);
        return CLSPMV_E_TOO_LARGE;
    }
    int rc = clspmv_alloc_coo(m, width, height, (int)start[height]);
    if (rc) return rc;
#pragma omp parallel num_threads(omp_threads())
    {
        std::vector<int> cols;
        std::vector<float> vals;
#pragma omp for schedule(dynamic, 1024)
        for (int r = 0; r < height; ++r) {
            cols.clear();
            vals.clear();
            fn(r, cols, vals);
            size_t o = (size_t)start[r];
            for (size_t k = 0; k < cols.size(); ++k) {
                m->coo_row_id[o + k] = r;
                m->coo_col_id[o + k] = cols[k];
                m->coo_data[o + k] = vals[k];
            }
        }
    }
    return CLSPMV_OK;
}

// Truncated Zipf(s) on [1, kmax] by inverse CDF.
struct ZipfTable {
    std::vector<double> cdf;
    double mean = 0.0;
    ZipfTable(double s, int kmax) {
        cdf.resize((size_t)kmax);
        double z = 0.0, m = 0.0;
        for (int k = 1; k <= kmax; ++k) {
            double p = std::pow((double)k, -s);
            z += p;
            m += p * k;
            cdf[(size_t)k - 1] = z;
        }
        for (auto& c : cdf) c /= z;
        mean = m / z;
    }
    int sample(Xoshiro& g) const {
        double u = g.uniform();
        return (int)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
    }
};

// Row lengths: k ~ truncated Zipf(s) on [1, max_len], rescaled to the target
// mean by L = max(1, round(k * mean / E[k])), capped at max_len; row
// (seed mod rows) is forced to max_len so the longest row is always present.
std::vector<int> zipf_lengths(int rows, double s, double mean, int max_len, uint64_t seed) {
    ZipfTable zt(s, max_len);
    const double scale = mean / zt.mean;
    std::vector<int> len((size_t)rows);
#pragma omp parallel for schedule(static) num_threads(omp_threads())
    for (int r = 0; r < rows; ++r) {
        Xoshiro g(seed ^ 0xA5A5A5A5ull, (uint64_t)r);
        int k = zt.sample(g);
        long L = std::lround(k * scale);
        len[r] = (int)std::min<long>(max_len, std::max<long>(1, L));
    }
    if (rows > 0) len[(size_t)(seed % (uint64_t)rows)] = max_len;
    return len;
}

}  // namespace

extern "C" {

int clspmv_gen_vector(float* x, int n, uint64_t seed) {
#pragma omp parallel for schedule(static) num_threads(omp_threads())
    for (int blk = 0; blk < (n + 4095) / 4096; ++blk) {
        Xoshiro g(seed ^ 0x7F4A7C15ull, (uint64_t)blk);
        for (int i = blk * 4096; i < std::min(n, blk * 4096 + 4096); ++i) x[i] = (float)(2.0 * g.uniform() - 1.0);
    }
    return CLSPMV_OK;
}

// 2D 5-point Laplacian on an nx x ny grid: 4 on the diagonal, -1 for each
// in-grid neighbour (row r = iy*nx + ix).
int clspmv_gen_laplacian2d(clspmv_coo_matrix* m, int nx, int ny) {
    if (nx <= 0 || ny <= 0) return CLSPMV_E_INVALID;
    const int n = nx * ny;
    std::vector<int> len((size_t)n);
    for (int r = 0; r < n; ++r) {
        int ix = r % nx, iy = r / nx;
        len[r] = 1 + (ix > 0) + (ix < nx - 1) + (iy > 0) + (iy < ny - 1);
    }
    return build_rows(m, n, n, len, [&](int r, std::vector<int>& c, std::vector<float>& v) {
        int ix = r % nx, iy = r / nx;
        if (iy > 0) { c.push_back(r - nx); v.push_back(-1.0f); }
        if (ix > 0) { c.push_back(r - 1); v.push_back(-1.0f); }
        c.push_back(r); v.push_back(4.0f);
        if (ix < nx - 1) { c.push_back(r + 1); v.push_back(-1.0f); }
        if (iy < ny - 1) { c.push_back(r + nx); v.push_back(-1.0f); }
    });
}

// 3D 27-point stencil on an n^3 grid, values N(0,1) (row r = (z*n + y)*n + x).
int clspmv_gen_stencil27(clspmv_coo_matrix* m, int n, uint64_t seed) {
    if (n <= 0 || (int64_t)n * n * n > INT32_MAX) return CLSPMV_E_INVALID;
    const int N = n * n * n;
    std::vector<int> len((size_t)N);
    auto cnt = [n](int i) { return 1 + (i > 0) + (i < n - 1); };
    for (int r = 0; r < N; ++r) len[r] = cnt(r % n) * cnt((r / n) % n) * cnt(r / (n * n));
    return build_rows(m, N, N, len, [&](int r, std::vector<int>& c, std::vector<float>& v) {
        Xoshiro g(seed, (uint64_t)r);
        int x = r % n, y = (r / n) % n, z = r / (n * n);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int xx = x + dx, yy = y + dy, zz = z + dz;
                    if (xx < 0 || yy < 0 || zz < 0 || xx >= n || yy >= n || zz >= n) continue;
                    c.push_back((zz * n + yy) * n + xx);
                    v.push_back(g.normal());
                }
    });
}

// `rows` x `rows`, exactly per_row distinct uniform columns per row, N(0,1).
int clspmv_gen_uniform_rows(clspmv_coo_matrix* m, int rows, int per_row, uint64_t seed) {
    if (rows <= 0 || per_row < 0 || per_row > rows) return CLSPMV_E_INVALID;
    std::vector<int> len((size_t)rows, per_row);
    return build_rows(m, rows, rows, len, [&](int r, std::vector<int>& c, std::vector<float>& v) {
        Xoshiro g(seed, (uint64_t)r);
        distinct_sorted(g, 0, rows, per_row, c);
        for (size_t k = 0; k < c.size(); ++k) v.push_back(g.normal());
    });
}

// `rows` x `rows`, Zipf(exponent) row lengths rescaled to `mean`, capped at
// max_len (one row forced to max_len), uniform distinct columns, N(0,1).
int clspmv_gen_zipf_rows(clspmv_coo_matrix* m, int rows, double exponent, double mean, int max_len, uint64_t seed) {
    if (rows <= 0 || max_len <= 0 || max_len > rows || exponent <= 0) return CLSPMV_E_INVALID;
    std::vector<int> len = zipf_lengths(rows, exponent, mean, max_len, seed);
    return build_rows(m, rows, rows, len, [&](int r, std::vector<int>& c, std::vector<float>& v) {
        Xoshiro g(seed, (uint64_t)r);
        distinct_sorted(g, 0, rows, len[r], c);
        for (size_t k = 0; k < c.size(); ++k) v.push_back(g.normal());
    });
}

// block_rows block rows of dense bdim x bdim blocks; each block row has
// blocks_per_row distinct block columns drawn uniformly from
// [i - band, i + band] clipped to the matrix; N(0,1) values.
int clspmv_gen_block_rows(clspmv_coo_matrix* m, int block_rows, int bdim, int blocks_per_row, int band,
                          uint64_t seed) {
    if (block_rows <= 0 || bdim <= 0 || blocks_per_row <= 0 || band < 0) return CLSPMV_E_INVALID;
    const int64_t n64 = (int64_t)block_rows * bdim;
    if (n64 > INT32_MAX) return CLSPMV_E_INVALID;
    const int n = (int)n64;
    // block columns per block row (shared by its bdim rows)
    std::vector<int> bcols((size_t)block_rows * blocks_per_row);
    std::vector<int> nbc((size_t)block_rows);
#pragma omp parallel num_threads(omp_threads())
    {
        std::vector<int> tmp;
#pragma omp for schedule(static)
        for (int i = 0; i < block_rows; ++i) {
            Xoshiro g(seed ^ 0xB10Cull, (uint64_t)i);
            int lo = std::max(0, i - band), hi = std::min(block_rows, i + band + 1);
            distinct_sorted(g, lo, hi, blocks_per_row, tmp);
            nbc[i] = (int)tmp.size();
            std::copy(tmp.begin(), tmp.end(), bcols.begin() + (size_t)i * blocks_per_row);
        }
    }
    std::vector<int> len((size_t)n);
    for (int r = 0; r < n; ++r) len[r] = nbc[r / bdim] * bdim;
    return build_rows(m, n, n, len, [&](int r, std::vector<int>& c, std::vector<float>& v) {
        Xoshiro g(seed, (uint64_t)r);
        int i = r / bdim;
        for (int b = 0; b < nbc[i]; ++b) {
            int bc = bcols[(size_t)i * blocks_per_row + b];
            for (int w = 0; w < bdim; ++w) {
                c.push_back(bc * bdim + w);
                v.push_back(g.normal());
            }
        }
    });
}

// Cocktail composite (rows multiple of 16):
//   [0, R/4)     7-diagonal band (offsets -3..3) + 2 scattered columns per row
//   [R/4, R/2)   dense 4x4 blocks, 6 distinct block columns per block row
//                within +-4096 block columns of the diagonal
//   [R/2, R)     Zipf(1.6) row lengths, mean 12, at most 1000, uniform columns
int clspmv_gen_cocktail_composite(clspmv_coo_matrix* m, int rows, uint64_t seed) {
    if (rows <= 0 || rows % 16 != 0) return CLSPMV_E_INVALID;
    const int q1 = rows / 4, q2 = rows / 2;
    const int zrows = rows - q2;
    std::vector<int> zlen = zipf_lengths(zrows, 1.6, 12.0, std::min(1000, rows), seed ^ 0x2);
    const int nbr = (q2 - q1) / 4;
    const int nbcol = rows / 4;
    std::vector<int> bcols((size_t)nbr * 6);
#pragma omp parallel num_threads(omp_threads())
    {
        std::vector<int> tmp;
#pragma omp for schedule(static)
        for (int i = 0; i < nbr; ++i) {
            Xoshiro g(seed ^ 0xB10Cull, (uint64_t)i);
            int diag = (q1 / 4) + i;
            int lo = std::max(0, diag - 4096), hi = std::min(nbcol, diag + 4096 + 1);
            distinct_sorted(g, lo, hi, 6, tmp);
            std::copy(tmp.begin(), tmp.end(), bcols.begin() + (size_t)i * 6);
        }
    }
    std::vector<int> len((size_t)rows);
    for (int r = 0; r < rows; ++r) {
        if (r < q1) {
            int band = 0;
            for (int d = -3; d <= 3; ++d) band += (r + d >= 0 && r + d < rows);
            len[r] = band + 2;
        } else if (r < q2) {
            len[r] = 24;
        } else {
            len[r] = zlen[(size_t)(r - q2)];
        }
    }
    return build_rows(m, rows, rows, len, [&](int r, std::vector<int>& c, std::vector<float>& v) {
        Xoshiro g(seed, (uint64_t)r);
        if (r < q1) {
            // two scattered columns outside the band
            int extra[2];
            int got = 0;
            while (got < 2) {
                int cc = (int)g.below((uint32_t)rows);
                if (cc >= r - 3 && cc <= r + 3) continue;
                if (got == 1 && cc == extra[0]) continue;
                extra[got++] = cc;
            }
            std::vector<int> all;
            for (int d = -3; d <= 3; ++d)
                if (r + d >= 0 && r + d < rows) all.push_back(r + d);
            all.push_back(extra[0]);
            all.push_back(extra[1]);
            std::sort(all.begin(), all.end());
            for (int cc : all) {
                c.push_back(cc);
                v.push_back(g.normal());
            }
        } else if (r < q2) {
            int i = (r - q1) / 4;
            for (int b = 0; b < 6; ++b)
                for (int w = 0; w < 4; ++w) {
                    c.push_back(bcols[(size_t)i * 6 + b] * 4 + w);
                    v.push_back(g.normal());
                }
        } else {
            distinct_sorted(g, 0, rows, len[r], c);
            for (size_t k = 0; k < c.size(); ++k) v.push_back(g.normal());
        }
    });
}

}  // extern "C"
