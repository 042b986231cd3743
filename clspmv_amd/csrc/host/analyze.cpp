// Cocktail analyzer: splits one matrix into up to nine sub-matrices, one per
// format, using per-format performance tables.
//
// This is a restatement of the reference's analyze_matrix
// (A/src/analyze.cpp:16-3900) written for this engine.  The decision
// arithmetic (float/double conversions, integer truncations, tie-breaking,
// the order in which the per-format maxima are recomputed) follows the
// reference step for step, because the partition must be bit-exact with the
// reference's for the same tables; the data handling is this engine's own
// (per-block-row sorted runs instead of std::map counting, hash lookups
// instead of linear diagonal searches).  Function comments cite the
// reference function each one restates.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "clspmv.h"
#include "clspmv_cocktail.h"
#include "internal.hpp"

using namespace clspmv;

namespace {

// Reference compile-time constants (A/include/constant.h).
constexpr int kBdiaImp = 5, kDiaImp = 4, kSbellImp = 1, kBellImp = 2, kBcsrImp = 1, kSellImp = 2, kEllImp = 2,
              kCsrImp = 2, kCooImp = 1;
constexpr int kWorkGroup = 256, kSellGroup = 128, kBellGroup = 256, kCsrVecGroup = 64, kCooGroup = 64;
constexpr int kWarp = 32, kGpuAlign = 32, kMaxDiaNum = 128;
constexpr int BH[8] = {1, 2, 4, 8, 1, 2, 4, 8};
constexpr int BW[8] = {4, 4, 4, 4, 8, 8, 8, 8};

template <int N>
struct FlatEnt {
    int dim, width;
    double g[N];
};
template <int N>
struct GroupEnt {
    int dim, width, groups;
    double g[N];
};
template <int N>
struct BlockEnt {
    int bh, bw, dim, blocks;
    double g[N];
};
struct OverEnt {
    int groups, threads;
    double us;
};

}  // namespace

struct clspmv_bench_s {
    std::vector<OverEnt> over;
    std::vector<FlatEnt<kBdiaImp>> bdia;
    std::vector<FlatEnt<kDiaImp>> dia;
    std::vector<BlockEnt<kSbellImp>> sbell;
    std::vector<BlockEnt<kBellImp>> bell;
    std::vector<BlockEnt<kBcsrImp>> bcsr;
    std::vector<FlatEnt<kSellImp>> sell;
    std::vector<FlatEnt<kEllImp>> ell;
    std::vector<GroupEnt<kCsrImp>> csr;
    std::vector<GroupEnt<kCooImp>> coo;
};

namespace {

// ------------------------------------------------------------ tables (A1)

bool read_numbers(const std::string& path, std::vector<double>& out) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    double v;
    while (std::fscanf(f, "%lf", &v) == 1) out.push_back(v);
    std::fclose(f);
    return true;
}

// Each record is `nint` integers followed by `ndbl` doubles, whitespace
// separated, as the reference's fscanf loops read them (analyze.cpp:16-175).
template <class F>
bool read_records(const std::string& path, int nint, int ndbl, F push) {
    std::vector<double> v;
    if (!read_numbers(path, v)) return false;
    const size_t rec = (size_t)(nint + ndbl);
    for (size_t i = 0; i + rec <= v.size(); i += rec) push(&v[i]);
    return true;
}

// ------------------------------------------------------------ interpolation (A2)

struct Four {
    int ss = 0, sl = 0, ls = 0, ll = 0;
};

// find_fourid (analyze.cpp:294-399): the four table entries bracketing
// (dim, width); entries are grouped by dim, widths ascending inside a group.
Four find_fourid(const std::vector<std::pair<int, int>>& b, int dim, int width) {
    std::vector<int> dims{b[0].first}, off{0};
    for (int i = 0; i < (int)b.size(); ++i)
        if (b[i].first != dims.back()) {
            dims.push_back(b[i].first);
            off.push_back(i);
        }
    off.push_back((int)b.size());
    int ds = 0, dl = 0;
    if (dim < dims[0]) {
        ds = dl = 0;
    } else if (dim > dims.back()) {
        ds = dl = (int)dims.size() - 1;
    } else {
        for (int i = 0; i < (int)dims.size(); ++i) {
            if (dim == dims[i]) {
                ds = dl = i;
                break;
            }
            if (i < (int)dims.size() - 1 && dims[i] < dim && dims[i + 1] > dim) {
                ds = i;
                dl = i + 1;
                break;
            }
        }
    }
    Four f;
    auto widths = [&](int d, int& s, int& l) {
        const int a = off[d], e = off[d + 1];
        if (width < b[a].second) {
            s = l = a;
        } else if (width > b[e - 1].second) {
            s = l = e - 1;
        } else {
            for (int i = a; i < e; ++i) {
                if (width == b[i].second) {
                    s = l = i;
                    break;
                }
                // no break: the last bracketing pair wins, as in the reference
                if (i < e - 1 && b[i].second < width && b[i + 1].second > width) {
                    s = i;
                    l = i + 1;
                }
            }
        }
    };
    widths(ds, f.ss, f.sl);
    widths(dl, f.ls, f.ll);
    return f;
}

// linear_inter (analyze.cpp:401-411).
double lin(int small, double sv, int large, double lv, int index) {
    if (small == large) return lv;
    return sv + ((double)(index - small)) * (lv - sv) / ((double)(large - small));
}

// Bilinear value of column k for a flat table (dimension x width).
template <class E, class G>
double bilinear(const std::vector<E>& t, const Four& f, G val, int dim, int width) {
    double sv = lin(t[f.ss].width, val(t[f.ss]), t[f.sl].width, val(t[f.sl]), width);
    double lv = lin(t[f.ls].width, val(t[f.ls]), t[f.ll].width, val(t[f.ll]), width);
    return lin(t[f.ss].dim, sv, t[f.ll].dim, lv, dim);
}

// maxflop_{bdia,dia,sell,ell} (analyze.cpp:431-541).
template <int N>
void maxflop_flat(const std::vector<FlatEnt<N>>& t, int rows, int npr, int& meth, double& flop) {
    std::vector<std::pair<int, int>> b(t.size());
    for (size_t i = 0; i < t.size(); ++i) b[i] = {t[i].dim, t[i].width};
    Four f = find_fourid(b, rows, npr);
    meth = 0;
    flop = 0.0;
    for (int k = 0; k < N; ++k) {
        double v = bilinear(t, f, [k](const FlatEnt<N>& e) { return e.g[k]; }, rows, npr);
        if (v > flop) {
            meth = k;
            flop = v;
        }
    }
}

// maxflop_{sbell,bell,bcsr} (analyze.cpp:543-706): per block shape, the
// table width is blocks per block row = round(nnz/row / bw).
template <int N>
void maxflop_block(const std::vector<BlockEnt<N>>& t, int rows, int npr, int meth[8], double flop[8]) {
    int off[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int b = 1;
    for (int i = 0; i < (int)t.size() && b < 8; ++i)
        if (t[i].bw == BW[b] && t[i].bh == BH[b]) off[b++] = i;
    off[8] = (int)t.size();
    for (b = 0; b < 8; ++b) {
        const int bn = (int)(((double)npr) / ((double)BW[b]) + 0.5);
        std::vector<std::pair<int, int>> basic;
        for (int i = off[b]; i < off[b + 1]; ++i) basic.push_back({t[i].dim, t[i].blocks});
        Four f = find_fourid(basic, rows, bn);
        f.ss += off[b];
        f.sl += off[b];
        f.ls += off[b];
        f.ll += off[b];
        meth[b] = 0;
        flop[b] = 0.0;
        for (int k = 0; k < N; ++k) {
            double sv = lin(t[f.ss].blocks, t[f.ss].g[k], t[f.sl].blocks, t[f.sl].g[k], bn);
            double lv = lin(t[f.ls].blocks, t[f.ls].g[k], t[f.ll].blocks, t[f.ll].g[k], bn);
            double v = lin(t[f.ss].dim, sv, t[f.ll].dim, lv, rows);
            if (v > flop[b]) {
                meth[b] = k;
                flop[b] = v;
            }
        }
    }
}

// maxflop_{csr,coo} (analyze.cpp:708-802): tables carry one row per
// workgroup count for every (dimension, width); the best (method, groups)
// pair wins.
template <int N>
void maxflop_group(const std::vector<GroupEnt<N>>& t, int rows, int npr, int& meth, int& groups, double& flop) {
    std::vector<int> off{0};
    int pd = t[0].dim, pw = t[0].width;
    for (int i = 0; i < (int)t.size(); ++i)
        if (pd != t[i].dim || pw != t[i].width) {
            off.push_back(i);
            pd = t[i].dim;
            pw = t[i].width;
        }
    off.push_back((int)t.size());
    std::vector<std::pair<int, int>> basic(off.size() - 1);
    for (size_t i = 0; i + 1 < off.size(); ++i) basic[i] = {t[off[i]].dim, t[off[i]].width};
    Four f = find_fourid(basic, rows, npr);
    meth = 0;
    flop = 0.0;
    const int choices = off[1] - off[0];
    for (int g = 0; g < choices; ++g)
        for (int k = 0; k < N; ++k) {
            const auto& ss = t[off[f.ss] + g];
            const auto& sl = t[off[f.sl] + g];
            const auto& ls = t[off[f.ls] + g];
            const auto& ll = t[off[f.ll] + g];
            double sv = lin(ss.width, ss.g[k], sl.width, sl.g[k], npr);
            double lv = lin(ls.width, ls.g[k], ll.width, ll.g[k], npr);
            double v = lin(ss.dim, sv, ll.dim, lv, rows);
            if (v > flop) {
                meth = k;
                groups = ss.groups;
                flop = v;
            }
        }
}

// find_overhead (analyze.cpp:413-428), microseconds.
double overhead_us(const clspmv_bench_s& B, int groups, int threads) {
    std::vector<std::pair<int, int>> b(B.over.size());
    for (size_t i = 0; i < B.over.size(); ++i) b[i] = {B.over[i].groups, B.over[i].threads};
    Four f = find_fourid(b, groups, threads);
    const auto& o = B.over;
    double sv = lin(o[f.ss].threads, o[f.ss].us, o[f.sl].threads, o[f.sl].us, threads);
    double lv = lin(o[f.ls].threads, o[f.ls].us, o[f.ll].threads, o[f.ll].us, threads);
    return lin(o[f.ss].groups, sv, o[f.ll].groups, lv, groups);
}
double overhead_ms(const clspmv_bench_s& B, int groups, int threads) { return overhead_us(B, groups, threads) / 1000.0; }

// nnz per row as the reference computes it everywhere: in float, +0.5, cast.
int nnz_per_row(int nnz, int rows) { return (int)(((float)nnz) / ((float)rows) + 0.5); }

// A part whose slot count does not fit the int32 slice pointers of its
// format (the reference would overflow and write out of bounds there).
struct TooLarge {
    const char* what;
    explicit TooLarge(const char* w) : what(w) {}
};

// ------------------------------------------------------------ COO helper

struct Coo {
    clspmv_coo_matrix m{};
    bool made = false;  // the reference distinguishes "never created" (NULL)
    Coo() = default;
    Coo(const Coo&) = delete;
    ~Coo() { clspmv_free_coo(&m); }
    void alloc(int w, int h, int nnz) {
        clspmv_free_coo(&m);
        clspmv_alloc_coo(&m, w, h, nnz);
        made = true;
    }
};

void copy_coo(const clspmv_coo_matrix& src, clspmv_coo_matrix& dst) {
    clspmv_alloc_coo(&dst, src.matinfo.width, src.matinfo.height, src.matinfo.nnz);
    const size_t n = (size_t)src.matinfo.nnz;
    if (n) {
        std::memcpy(dst.coo_row_id, src.coo_row_id, n * 4);
        std::memcpy(dst.coo_col_id, src.coo_col_id, n * 4);
        std::memcpy(dst.coo_data, src.coo_data, n * 4);
    }
}

std::vector<int> row_counts(const clspmv_coo_matrix& m) {
    std::vector<int> c((size_t)m.matinfo.height, 0);
    for (int k = 0; k < m.matinfo.nnz; ++k) c[(size_t)m.coo_row_id[k]]++;
    return c;
}

std::vector<int> rowptr_of(const clspmv_coo_matrix& m) {
    std::vector<int> c = row_counts(m), rp((size_t)m.matinfo.height + 1, 0);
    for (size_t i = 0; i < c.size(); ++i) rp[i + 1] = rp[i] + c[i];
    return rp;
}

template <class T>
T* xmalloc(size_t n) {
    return static_cast<T*>(std::malloc(n ? n * sizeof(T) : 1));
}
template <class T>
T* xcalloc(size_t n) {
    return static_cast<T*>(std::calloc(n ? n : 1, sizeof(T)));
}

// ------------------------------------------------------------ analyzer

class Analyzer {
  public:
    Analyzer(const clspmv_bench_s& b, clspmv_cocktail& mat, bool verbose) : B(b), mat(mat), verbose(verbose) {}

    void run(clspmv_coo_matrix& coo, float dia_mult, bool est_partial, bool est_full);

  private:
    const clspmv_bench_s& B;
    clspmv_cocktail& mat;
    bool verbose;
    clspmv_maxflop_info info{};

    void log(const char* fmt, ...) {
        if (!verbose) return;
        va_list ap;
        va_start(ap, fmt);
        std::vfprintf(stderr, fmt, ap);
        va_end(ap);
    }

  public:
    // update_info_* (analyze.cpp:1311-1363)
    void update_dia(int rows, int npr) {
        maxflop_flat(B.bdia, rows, npr, info.bdia_meth_num, info.bdia_max_flop);
        maxflop_flat(B.dia, rows, npr, info.dia_meth_num, info.dia_max_flop);
        info.max_dia_base_flop = info.dia_max_flop > info.bdia_max_flop ? info.dia_max_flop : info.bdia_max_flop;
    }
    void update_block(int rows, int npr) {
        maxflop_block(B.sbell, rows, npr, info.sbell_meth_num, info.sbell_max_flop);
        maxflop_block(B.bell, rows, npr, info.bell_meth_num, info.bell_max_flop);
        maxflop_block(B.bcsr, rows, npr, info.bcsr_meth_num, info.bcsr_max_flop);
        double m = 0.0;
        for (int i = 0; i < 8; ++i) {
            if (info.sbell_max_flop[i] > m) m = info.sbell_max_flop[i];
            if (info.bell_max_flop[i] > m) m = info.bell_max_flop[i];
            if (info.bcsr_max_flop[i] > m) m = info.bcsr_max_flop[i];
        }
        info.max_block_base_flop = m;
    }
    static void update_flat(const clspmv_bench_s& B, clspmv_maxflop_info& in, int rows, int npr) {
        maxflop_flat(B.sell, rows, npr, in.sell_meth_num, in.sell_max_flop);
        maxflop_flat(B.ell, rows, npr, in.ell_meth_num, in.ell_max_flop);
        maxflop_group(B.csr, rows, npr, in.csr_meth_num, in.csr_groupnum, in.csr_max_flop);
        maxflop_group(B.coo, rows, npr, in.coo_meth_num, in.coo_groupnum, in.coo_max_flop);
        double m = in.sell_max_flop;
        if (in.ell_max_flop > m) m = in.ell_max_flop;
        if (in.csr_max_flop > m) m = in.csr_max_flop;
        if (in.coo_max_flop > m) m = in.coo_max_flop;
        in.max_flat_base_flop = m;
    }
    void update_flat(int rows, int npr) { update_flat(B, info, rows, npr); }
    clspmv_maxflop_info& state() { return info; }

  private:
    bool extract_dia(clspmv_coo_matrix& coo, Coo& rem);
    bool extract_block(clspmv_coo_matrix& coo, Coo& rem, bool est_partial, bool est_full);
    void extract_flat(clspmv_coo_matrix& coo);
    void extract_csr_coo(clspmv_coo_matrix& coo);
    void extract_coo(clspmv_coo_matrix& coo);
    void extract_csr(clspmv_coo_matrix& coo);
    void extract_sell_ell(clspmv_coo_matrix& coo, Coo& rem);
    void extract_sell(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& cnt, int S,
                      const std::vector<int>& slice_ell, int slice_nnz);
    void extract_ell(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& cnt, int ellnum, int ellnnz);
    void place_diagonals(clspmv_coo_matrix& coo, Coo& rem, const std::vector<std::pair<int, int>>& select,
                         int threshold, int kind, int least_band, int bdia_meth, int dia_meth);
    void extract_full_sbell_bell(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& rp, int wh,
                                 const std::vector<std::pair<int, int>>& bc, std::pair<int, int> sum);
    void extract_sbell_bell(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& rp, int wh,
                            const std::vector<std::pair<int, int>>& bc, std::pair<int, int> sum);
    void extract_sbell_part(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& rp, int wh, int threshold,
                            int meth, const std::vector<int>& slice_num_blocks, int S);
    void extract_bell_part(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& rp, int wh, int threshold,
                           int meth, int bell_num);
    void extract_bcsr_part(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& rp, int wh, int bcsr_nnz,
                           int threshold, int meth);

  public:
    double estimate_flat_time(const std::vector<int>& rownnz, int nnz);
    static double estimate_coo_time(const clspmv_bench_s& B, int nnz, const clspmv_maxflop_info& in);
    static double estimate_csr_time(const clspmv_bench_s& B, const std::vector<int>& rownnz, int nnz,
                                    const clspmv_maxflop_info& in);
    double estimate_csr_coo_time(const std::vector<int>& rownnz, int nnz, clspmv_maxflop_info& in);
    double estimate_sell_ell_time(const std::vector<int>& rownnz, int nnz, clspmv_maxflop_info& in,
                                  std::vector<int>& remain, int& nnzremain);
};

// ---- estimators (analyze.cpp:2846-2958, 3258-3449, 3629-3673) ----

double Analyzer::estimate_coo_time(const clspmv_bench_s& B, int nnz, const clspmv_maxflop_info& in) {
    if (nnz == 0) return 0.0;
    return ((double)nnz) / 1000000.0 / in.coo_max_flop + overhead_ms(B, in.coo_groupnum, kCooGroup);
}

// Worst workgroup load of the reference's CSR kernels: method 0 gives each
// 32-lane warp one row, strided over groups*2 warps; method 1 gives each
// workgroup one row, strided over the groups.
static int csr_max_group_nnz(const std::vector<int>& rownnz, int groups, int meth) {
    std::vector<int> per((size_t)std::max(0, groups), 0);
    const int n = (int)rownnz.size();
    if (meth == 0) {
        const int hgt = kCsrVecGroup / kWarp;
        const int warps = groups * hgt;
        int iters = n / warps + (n % warps != 0);
        for (int i = 0; i < iters; ++i)
            for (int g = 0; g < groups; ++g)
                for (int j = 0; j < hgt; ++j) {
                    int id = i * warps + g * hgt + j;
                    if (id >= n) break;
                    per[g] += rownnz[id];
                }
    } else if (meth == 1) {
        for (int i = 0; i < n; ++i)
            for (int g = 0; g < groups; ++g) {
                long long id = (long long)i * groups + g;
                if (id >= n) break;
                per[g] += rownnz[(size_t)id];
            }
    }
    int mx = 0;
    for (int v : per) mx = std::max(mx, v);
    return mx;
}

double Analyzer::estimate_csr_time(const clspmv_bench_s& B, const std::vector<int>& rownnz, int nnz,
                                   const clspmv_maxflop_info& in) {
    if (nnz == 0) return 0.0;
    int newnnz = csr_max_group_nnz(rownnz, in.csr_groupnum, in.csr_meth_num) * in.csr_groupnum;
    return ((double)newnnz) / 1000000.0 / in.csr_max_flop + overhead_ms(B, in.csr_groupnum, kCsrVecGroup);
}

double Analyzer::estimate_csr_coo_time(const std::vector<int>& rownnz, int nnz, clspmv_maxflop_info& in) {
    if (nnz == 0) return 0.0;
    const int npr = nnz_per_row(nnz, (int)rownnz.size());
    maxflop_group(B.csr, (int)rownnz.size(), npr, in.csr_meth_num, in.csr_groupnum, in.csr_max_flop);
    maxflop_group(B.coo, (int)rownnz.size(), npr, in.coo_meth_num, in.coo_groupnum, in.coo_max_flop);
    if (in.coo_max_flop >= in.csr_max_flop) return estimate_coo_time(B, nnz, in);
    double ct = estimate_coo_time(B, nnz, in), rt = estimate_csr_time(B, rownnz, nnz, in);
    return ct < rt ? ct : rt;
}

// ELL width and SELL slice widths that balance padding against the flat
// remainder (shared by estimate_sell_ell_time and extract_sell_ell).
struct SellEllPlan {
    double ell_time, sell_full_time, sell_partial_time;
    int ellnum, ellnnz, S;
    std::vector<int> slice_partial, slice_full;
    int slice_partial_nnz, slice_full_nnz;
};

static SellEllPlan plan_sell_ell(const clspmv_bench_s& B, const clspmv_maxflop_info& in, const std::vector<int>& rownnz,
                                 int nnz, int height) {
    SellEllPlan p{};
    double rflop;
    int rgroup, rthread;
    if (in.csr_max_flop > in.coo_max_flop) {
        rflop = in.csr_max_flop;
        rgroup = in.csr_groupnum;
        rthread = kCsrVecGroup;
    } else {
        rflop = in.coo_max_flop;
        rgroup = in.coo_groupnum;
        rthread = kCooGroup;
    }
    const double rover = overhead_ms(B, rgroup, rthread);
    int maxr = 0;
    for (int v : rownnz) maxr = std::max(maxr, v);
    std::vector<int> hist((size_t)maxr + 1, 0);
    for (int v : rownnz) hist[(size_t)v]++;
    int rownum = height;
    int ellnnz = 0;
    double best = 1000000.0;
    int bestnum = 0, bestnnz = 0;
    double eover = 0.0;
    if (in.ell_meth_num == 0)
        eover = overhead_us(B, height / kWorkGroup, kWorkGroup);
    else if (in.ell_meth_num == 1)
        eover = overhead_us(B, height / (4 * kWorkGroup), kWorkGroup);
    eover /= 1000.0;
    const int newlength = clspmv_aligned_length(height, kGpuAlign);
    for (int i = 0; i < (int)hist.size(); ++i) {
        rownum -= hist[(size_t)i];
        double t = 0.0;
        if (i > 0) t += ((double)i * (double)newlength) / 1000000.0 / in.ell_max_flop + eover;
        if (ellnnz < nnz) t += ((double)(nnz - ellnnz)) / 1000000.0 / rflop + rover;
        if (t < best) {
            best = t;
            bestnum = i;
            bestnnz = ellnnz;
        }
        ellnnz += rownum;
    }
    p.ell_time = best;
    p.ellnum = bestnum;
    p.ellnnz = bestnnz;
    int S = 0;
    if (in.sell_meth_num == 0)
        S = kWarp;
    else if (in.sell_meth_num == 1)
        S = kSellGroup;
    p.S = S;
    const int nrow = (int)rownnz.size();
    const int ns = nrow / S + (nrow % S != 0);
    p.slice_partial.assign((size_t)ns, 0);
    p.slice_full.assign((size_t)ns, 0);
    p.slice_partial_nnz = 0;
    p.slice_full_nnz = nnz;
    const double ratio = rflop / in.sell_max_flop;
    std::vector<int> w;
    for (int i = 0; i < ns; ++i) {
        w.clear();
        for (int j = i * S; j < (i + 1) * S && j < nrow; ++j) w.push_back(rownnz[(size_t)j]);
        std::sort(w.begin(), w.end());
        p.slice_full[(size_t)i] = w.back();
        const size_t n = w.size();
        int partial = (int)((size_t)w[0] * n);
        int current = partial;
        p.slice_partial[(size_t)i] = w[0];
        for (size_t j = 1; j < n; ++j) {
            partial += (int)((size_t)(w[j] - w[j - 1]) * (n - j));
            if (partial > (int)(((double)(n * (size_t)w[j])) * ratio)) {
                p.slice_partial[(size_t)i] = w[j];
                current = partial;
            } else {
                break;
            }
        }
        p.slice_partial_nnz += current;
    }
    const double sover = overhead_ms(B, ns, kSellGroup);
    // Slot totals in 64 bits: the reference sums them in int, which
    // overflows (undefined) for matrices with very long rows; wherever it
    // does not overflow the value is the same.
    long long full = 0, part = 0;
    for (int v : p.slice_full) full += (long long)S * v;
    for (int v : p.slice_partial) part += (long long)S * v;
    p.sell_full_time = ((double)full) / 1000000.0 / in.sell_max_flop + sover;
    p.sell_partial_time = ((double)part) / 1000000.0 / in.sell_max_flop + sover +
                          ((double)(nnz - p.slice_partial_nnz)) / 1000000.0 / rflop + rover;
    return p;
}

double Analyzer::estimate_sell_ell_time(const std::vector<int>& rownnz, int nnz, clspmv_maxflop_info& in,
                                        std::vector<int>& remain, int& nnzremain) {
    if (nnz == 0) {
        nnzremain = 0;
        std::fill(remain.begin(), remain.end(), 0);
        return 0.0;
    }
    SellEllPlan p = plan_sell_ell(B, in, rownnz, nnz, (int)rownnz.size());
    nnzremain = 0;
    if (p.ell_time < p.sell_full_time && p.ell_time < p.sell_partial_time) {
        for (size_t i = 0; i < rownnz.size(); ++i) {
            int d = rownnz[i] > p.ellnum ? rownnz[i] - p.ellnum : 0;
            remain[i] = d;
            nnzremain += d;
        }
        return p.ell_time;
    }
    if (p.sell_partial_time < p.sell_full_time) {
        for (size_t i = 0; i < rownnz.size(); ++i) {
            int wdt = p.slice_partial[i / (size_t)p.S];
            int d = rownnz[i] > wdt ? rownnz[i] - wdt : 0;
            remain[i] = d;
            nnzremain += d;
        }
        return p.sell_partial_time;
    }
    std::fill(remain.begin(), remain.end(), 0);
    return p.sell_full_time;
}

double Analyzer::estimate_flat_time(const std::vector<int>& rownnz, int nnz) {
    if (nnz == 0) return 0.0;
    clspmv_maxflop_info in{};
    update_flat(B, in, (int)rownnz.size(), nnz_per_row(nnz, (int)rownnz.size()));
    if (in.max_flat_base_flop == in.coo_max_flop) return estimate_coo_time(B, nnz, in);
    if (in.max_flat_base_flop == in.csr_max_flop) return estimate_csr_time(B, rownnz, nnz, in);
    std::vector<int> remain(rownnz.size(), 0);
    int nremain = 0;
    double t = estimate_sell_ell_time(rownnz, nnz, in, remain, nremain);
    return t + estimate_csr_coo_time(remain, nremain, in);
}

// ---- diagonals (analyze.cpp:804-1524) ----

struct BdiaVsDia {
    int min_width = 1000000, range_s = 0, range_l = 0;
    double bdia_single = 0.0, dia_single = 0.0;
};

// find_min_bdia_width (analyze.cpp:828-911): in the table block of the
// dimension nearest to `dimension`, the first band width at which BDIA beats
// DIA.
BdiaVsDia find_min_bdia_width(const clspmv_bench_s& B, int dimension) {
    const auto& bd = B.bdia;
    const auto& di = B.dia;
    std::vector<int> dims{bd[0].dim}, off{0};
    for (int i = 0; i < (int)bd.size(); ++i)
        if (bd[i].dim != dims.back()) {
            dims.push_back(bd[i].dim);
            off.push_back(i);
        }
    off.push_back((int)bd.size());
    int ds = 0, dl = 0;
    if (dimension < dims[0])
        ds = dl = 0;
    else if (dimension > dims.back())
        ds = dl = (int)dims.size() - 1;
    else
        for (int i = 0; i < (int)dims.size(); ++i) {
            if (dimension == dims[i]) {
                ds = dl = i;
                break;
            }
            if (i < (int)dims.size() - 1 && dims[i] < dimension && dims[i + 1] > dimension) {
                ds = i;
                dl = i + 1;
                break;
            }
        }
    int near = ds;
    if (ds != dl) {
        int dist_s = dimension - bd[(size_t)off[ds]].dim;
        int dist_l = bd[(size_t)off[dl]].dim - dimension;
        near = (dist_s < dist_l) ? ds : dl;
    }
    BdiaVsDia r;
    r.range_s = off[near];
    r.range_l = off[near + 1];
    for (int i = off[near]; i < off[near + 1]; ++i) {
        double mb = 0.0, md = 0.0;
        for (int j = 0; j < kBdiaImp; ++j)
            if (bd[i].g[j] > mb) mb = bd[i].g[j];
        for (int j = 0; j < kDiaImp; ++j)
            if (di[i].g[j] > md) md = di[i].g[j];
        if (bd[i].width == 1) {
            r.bdia_single = mb;
            r.dia_single = md;
        }
        if (mb > md) {
            r.min_width = bd[i].width;
            return r;
        }
    }
    r.min_width = 1000000;
    return r;
}

// bdia_gflop / dia_gflop (analyze.cpp:913-987): best method at the table
// entry whose width brackets `width` from below.
template <int N>
void table_gflop(const std::vector<FlatEnt<N>>& t, int rs, int rl, int width, double& g, int& meth) {
    int idx = 0;
    if (width < t[rs].width)
        idx = rs;
    else if (width > t[rl - 1].width)
        idx = rl - 1;
    else
        for (int i = rs; i < rl; ++i) {
            if (width == t[i].width) {
                idx = i;
                break;
            }
            if (i < rl - 1 && t[i].width < width && t[i + 1].width > width) {
                idx = i;
                break;
            }
        }
    g = 0.0;
    meth = 0;
    for (int k = 0; k < N; ++k)
        if (t[idx].g[k] > g) {
            g = t[idx].g[k];
            meth = k;
        }
}

// extract_bdia_only / extract_dia_only / extract_bdia_and_dia
// (analyze.cpp:989-1309).  kind 0: BDIA only, 1: DIA only, 2: bands of at
// least `least_band` diagonals to BDIA and the rest to DIA.
void Analyzer::place_diagonals(clspmv_coo_matrix& coo, Coo& rem, const std::vector<std::pair<int, int>>& select,
                               int threshold, int kind, int least_band, int bdia_meth, int dia_meth) {
    const int h = coo.matinfo.height, w = coo.matinfo.width;
    const int L = clspmv_aligned_length(h, kGpuAlign);
    std::vector<int> off(select.size());
    std::vector<int> cnt(select.size());
    int total = 0;
    for (size_t i = 0; i < select.size(); ++i) {
        off[i] = select[i].first - threshold;
        cnt[i] = select[i].second;
        total += cnt[i];
    }
    // runs of consecutive offsets
    std::vector<int> band_off, band_cnt, band_nnz;
    for (size_t i = 0; i < off.size(); ++i) {
        if (i > 0 && off[i] - off[i - 1] == 1) {
            band_cnt.back()++;
            band_nnz.back() += cnt[i];
        } else {
            band_off.push_back(off[i]);
            band_cnt.push_back(1);
            band_nnz.push_back(cnt[i]);
        }
    }
    std::vector<int> dia_off, bdia_off;  // offsets in storage order
    std::vector<int> bands_o, bands_c;
    int dia_nnz = 0, bdia_nnz = 0;
    for (size_t b = 0; b < band_off.size(); ++b) {
        bool to_bdia = (kind == 0) || (kind == 2 && (unsigned)band_cnt[b] >= (unsigned)least_band);
        for (int j = 0; j < band_cnt[b]; ++j) (to_bdia ? bdia_off : dia_off).push_back(band_off[b] + j);
        if (to_bdia) {
            bands_o.push_back(band_off[b]);
            bands_c.push_back(band_cnt[b]);
            bdia_nnz += band_nnz[b];
        } else {
            dia_nnz += band_nnz[b];
        }
    }
    if (kind == 1) {  // DIA only: every selected diagonal, ascending
        dia_off = off;
        dia_nnz = total;
    }
    // kind 2 marks both parts used even when one of them is empty, as the
    // reference does (analyze.cpp:1149-1152)
    if (kind == 0 || kind == 2) {
        mat.ifusebdia = 1;
        mat.bdia_meth_num = bdia_meth;
        auto& b = mat.bdia;
        b.matinfo.width = w;
        b.matinfo.height = h;
        b.matinfo.nnz = bdia_nnz;
        b.bdia_length = h;
        b.bdia_length_aligned = L;
        b.bdia_band_num = (int)bands_o.size();
        b.bdia_offsets = xmalloc<int>(bands_o.size());
        b.bdia_bptr = xmalloc<int>(bands_o.size() + 1);
        b.bdia_bptr[0] = 0;
        for (size_t i = 0; i < bands_o.size(); ++i) {
            b.bdia_offsets[i] = bands_o[i];
            b.bdia_bptr[i + 1] = b.bdia_bptr[i] + bands_c[i];
        }
        b.bdia_data = xcalloc<float>(bdia_off.size() * (size_t)L);
    }
    if (kind == 1 || kind == 2) {
        mat.ifusedia = 1;
        mat.dia_meth_num = dia_meth;
        auto& d = mat.dia;
        d.matinfo.width = w;
        d.matinfo.height = h;
        d.matinfo.nnz = dia_nnz;
        d.dia_length = h;
        d.dia_length_aligned = L;
        d.dia_num = (int)dia_off.size();
        d.dia_offsets = xmalloc<int>(dia_off.size());
        for (size_t i = 0; i < dia_off.size(); ++i) d.dia_offsets[i] = dia_off[i];
        d.dia_data = xcalloc<float>(dia_off.size() * (size_t)L);
    }
    // place entries: DIA offsets are tested before BDIA offsets
    std::unordered_map<int, int> dslot, bslot;
    for (size_t i = 0; i < dia_off.size(); ++i) dslot.emplace(dia_off[i], (int)i);
    for (size_t i = 0; i < bdia_off.size(); ++i) bslot.emplace(bdia_off[i], (int)i);
    const int nremain = coo.matinfo.nnz - total;
    rem.alloc(w, h, nremain);
    int rid = 0;
    for (int k = 0; k < coo.matinfo.nnz; ++k) {
        const int r = coo.coo_row_id[k], c = coo.coo_col_id[k];
        const float v = coo.coo_data[k];
        const int o = c - r;
        auto it = dslot.find(o);
        if (it != dslot.end()) {
            mat.dia.dia_data[(size_t)it->second * L + r] = v;
            continue;
        }
        it = bslot.find(o);
        if (it != bslot.end()) {
            mat.bdia.bdia_data[(size_t)it->second * L + r] = v;
            continue;
        }
        rem.m.coo_row_id[rid] = r;
        rem.m.coo_col_id[rid] = c;
        rem.m.coo_data[rid] = v;
        ++rid;
    }
}

// extract_dia (analyze.cpp:1365-1524).
bool Analyzer::extract_dia(clspmv_coo_matrix& coo, Coo& rem) {
    const int h = coo.matinfo.height, nnz = coo.matinfo.nnz;
    if (nnz == 0) return false;
    const int npr = nnz_per_row(nnz, h);
    update_dia(h, npr);
    update_flat(h, npr);
    const double margin = 1.05;
    const double ratio = info.max_flat_base_flop / info.max_dia_base_flop;
    const int least = (int)(ratio * (double)h * margin);
    if (least > h) {
        log("No bdia or dia extracted\n");
        return false;
    }
    const int threshold = h - least;
    std::vector<int> diacount((size_t)(2 * threshold + 1), 0);
    for (int k = 0; k < nnz; ++k) {
        int o = coo.coo_col_id[k] - coo.coo_row_id[k] + threshold;
        if (o >= 0 && o < (int)diacount.size()) diacount[(size_t)o]++;
    }
    std::vector<std::pair<int, int>> valid;  // (count, index)
    for (int i = 0; i < (int)diacount.size(); ++i)
        if (diacount[(size_t)i] > least) valid.push_back({diacount[(size_t)i], i});
    std::sort(valid.begin(), valid.end(), std::greater<std::pair<int, int>>());
    const double over = overhead_ms(B, h / kWorkGroup, kWorkGroup);
    int cur = 0, ndia = 0;
    double best = 1000000.0;
    for (int i = 0; i <= kMaxDiaNum && i <= (int)valid.size(); ++i) {
        if (i > 0) cur += valid[(size_t)i - 1].first;
        const int remain = nnz - cur;
        double t = ((double)i) * ((double)h) / 1000000.0 / info.max_dia_base_flop +
                   ((double)remain) / 1000000.0 / info.max_flat_base_flop;
        if (i > 0) t += over;
        if (remain > 0) t += over;
        if (t < best) {
            best = t;
            ndia = i;
        }
    }
    log("max flat %f max dia %f overhead %f: extract %d diagonals\n", info.max_flat_base_flop,
        info.max_dia_base_flop, over, ndia);
    if (ndia == 0) return false;
    std::vector<std::pair<int, int>> select;  // (index, count)
    for (int i = 0; i < ndia; ++i) select.push_back({valid[(size_t)i].second, valid[(size_t)i].first});
    std::sort(select.begin(), select.end());
    const BdiaVsDia bv = find_min_bdia_width(B, h);
    const int least_band = bv.min_width;
    int band_num = 0, width = 1;
    for (size_t i = 1; i < select.size(); ++i) {
        if (select[i].first == select[i - 1].first + 1) {
            width++;
        } else {
            if (width >= least_band) band_num += width;
            width = 1;
        }
        if (i == select.size() - 1 && width >= least_band) band_num += width;
    }
    const int remain_dia = ndia - band_num;
    double bp = (double)0.0001f, dp = (double)0.0001f, da = (double)0.0001f;
    int bpm = 0, dpm = 0, dam = 0;
    if (band_num > 0) table_gflop(B.bdia, bv.range_s, bv.range_l, band_num, bp, bpm);
    if (remain_dia > 0) table_gflop(B.dia, bv.range_s, bv.range_l, remain_dia, dp, dpm);
    table_gflop(B.dia, bv.range_s, bv.range_l, ndia, da, dam);
    const double H = (double)h;
    const double bdia_only = ((double)band_num * H) / 1000000.0 / bp +
                             ((double)remain_dia * H) / 1000000.0 / (dp * bv.bdia_single / bv.dia_single) + over;
    const double dia_only = ((double)ndia * H) / 1000000.0 / da + over;
    const double both = ((double)band_num * H) / 1000000.0 / bp + ((double)remain_dia * H) / 1000000.0 / dp + 2 * over;
    if (bdia_only <= dia_only && bdia_only <= both) {
        log("Use bdia only\n");
        place_diagonals(coo, rem, select, threshold, 0, least_band, bpm, dpm);
    } else if (dia_only <= bdia_only && dia_only <= both) {
        log("Use dia only\n");
        place_diagonals(coo, rem, select, threshold, 1, least_band, bpm, dam);
    } else {
        log("Use bdia and dia\n");
        place_diagonals(coo, rem, select, threshold, 2, least_band, bpm, dpm);
    }
    return true;
}

// ---- blocks (analyze.cpp:1534-2844) ----

struct BlockInfo {
    int partial_block_num = 0, partial_nnz = 0, full_block_num = 0, full_nnz = 0;
};

// count_blocks (analyze.cpp:1582-1658): per shape and block row, the number
// of distinct blocks and their nnz, and those with more than threshold[i]
// entries.  Counted from each block row's sorted block columns.
void count_blocks(const clspmv_coo_matrix& m, const std::vector<int>& rp, const int thr[8],
                  std::vector<std::vector<BlockInfo>>& bc) {
    const int h = m.matinfo.height;
    bc.assign(8, {});
    for (int s = 0; s < 8; ++s) bc[(size_t)s].assign((size_t)(h / BH[s] + (h % BH[s] != 0)), BlockInfo{});
#pragma omp parallel for schedule(dynamic, 1) num_threads(omp_threads())
    for (int s = 0; s < 8; ++s) {
        std::vector<int> cols;
        for (int br = 0; br < (int)bc[(size_t)s].size(); ++br) {
            const int r0 = br * BH[s], r1 = std::min(h, r0 + BH[s]);
            cols.clear();
            for (int k = rp[(size_t)r0]; k < rp[(size_t)r1]; ++k) cols.push_back(m.coo_col_id[k] / BW[s]);
            std::sort(cols.begin(), cols.end());
            BlockInfo& b = bc[(size_t)s][(size_t)br];
            for (size_t i = 0; i < cols.size();) {
                size_t j = i;
                while (j < cols.size() && cols[j] == cols[i]) ++j;
                const int n = (int)(j - i);
                b.full_block_num++;
                b.full_nnz += n;
                if (n > thr[s]) {
                    b.partial_block_num++;
                    b.partial_nnz += n;
                }
                i = j;
            }
        }
    }
}

// Blocks of each block row, formed as the reference does (entries ordered by
// (col, row); consecutive entries with one block column form a block), and
// kept or returned to the remainder by `keep(block_row, count_if_kept,
// entries_in_block)`.
struct Picked {
    std::vector<int> rowptr, bcol;
    std::vector<float> data;
    std::vector<int> rr, rc;
    std::vector<float> rv;
};

struct El {
    int r, c;
    float v;
};

template <class Keep>
void pick_blocks(const clspmv_coo_matrix& m, const std::vector<int>& rp, int bw, int bh, Keep keep, Picked& P) {
    const int h = m.matinfo.height;
    const int brows = h / bh + (h % bh != 0);
    const int bs = bw * bh;
    P.rowptr.assign((size_t)brows + 1, 0);
    std::vector<El> el;
    for (int br = 0; br < brows; ++br) {
        const int r0 = br * bh;
        const int a = rp[(size_t)r0], e = rp[(size_t)std::min(h, r0 + bh)];
        int nb = 0;
        if (e > a) {
            el.resize((size_t)(e - a));
            for (int k = a; k < e; ++k) el[(size_t)(k - a)] = {m.coo_row_id[k], m.coo_col_id[k], m.coo_data[k]};
            std::sort(el.begin(), el.end(), [](const El& x, const El& y) {
                if (x.c != y.c) return x.c < y.c;
                return x.r < y.r;
            });
            size_t i = 0;
            while (i < el.size()) {
                const int bcol = el[i].c - el[i].c % bw;
                size_t j = i;
                while (j < el.size() && el[j].c - el[j].c % bw == bcol) ++j;
                if (keep(br, nb + 1, (int)(j - i))) {
                    ++nb;
                    const size_t base = P.data.size();
                    P.data.resize(base + (size_t)bs, 0.0f);
                    for (size_t q = i; q < j; ++q)
                        P.data[base + (size_t)((el[q].r % bh) * bw + (el[q].c - bcol))] = el[q].v;
                    P.bcol.push_back(bcol);
                } else {
                    for (size_t q = i; q < j; ++q) {
                        P.rr.push_back(el[q].r);
                        P.rc.push_back(el[q].c);
                        P.rv.push_back(el[q].v);
                    }
                }
                i = j;
            }
        }
        P.rowptr[(size_t)br + 1] = P.rowptr[(size_t)br] + nb;
    }
}

void remainder_from(const Picked& P, int w, int h, Coo& rem) {
    rem.alloc(w, h, (int)P.rr.size());
    if (!P.rr.empty()) {
        std::memcpy(rem.m.coo_row_id, P.rr.data(), P.rr.size() * 4);
        std::memcpy(rem.m.coo_col_id, P.rc.data(), P.rc.size() * 4);
        std::memcpy(rem.m.coo_data, P.rv.data(), P.rv.size() * 4);
    }
    if (!clspmv_if_sorted_coo(&rem.m)) clspmv_sort_coo(&rem.m);
}

// extract_sbell_part (analyze.cpp:1660-1886).
void Analyzer::extract_sbell_part(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& rp, int wh, int threshold,
                                  int meth, const std::vector<int>& sn, int S) {
    if (coo.matinfo.nnz == 0) return;
    const int bw = BW[wh], bh = BH[wh], bs = bw * bh;
    const int h = coo.matinfo.height;
    mat.ifusesbell = 1;
    mat.sbell_meth_num = meth;
    Picked P;
    pick_blocks(coo, rp, bw, bh,
                [&](int br, int nb, int n) { return n > threshold && nb <= sn[(size_t)(br / S)]; }, P);
    remainder_from(P, coo.matinfo.width, h, rem);
    auto& s = mat.sbell;
    s.matinfo.width = coo.matinfo.width;
    s.matinfo.height = h;
    s.matinfo.nnz = coo.matinfo.nnz - (int)P.rr.size();
    s.sbell_bwidth = (unsigned)bw;
    s.sbell_bheight = (unsigned)bh;
    const int brows = (int)P.rowptr.size() - 1;
    const int ns = (int)sn.size();
    s.sbell_slice_height = (unsigned)S;
    s.sbell_slice_num = ns;
    s.sbell_row_num = brows;
    long long slots = 0;
    for (int v : sn) slots += (long long)v * S;
    if (slots > INT32_MAX) throw TooLarge("SBELL part");
    s.sbell_slice_ptr = xmalloc<int>((size_t)ns + 1);
    s.sbell_slice_ptr[0] = 0;
    for (int i = 0; i < ns; ++i) s.sbell_slice_ptr[i + 1] = s.sbell_slice_ptr[i] + sn[(size_t)i] * S;
    const size_t total = (size_t)s.sbell_slice_ptr[ns];
    s.sbell_col_id = xcalloc<int>(total);
    s.sbell_data = xcalloc<float>(total * (size_t)bs);
    const int w4n = bw / 4;
    const size_t slot_stride = (size_t)S * w4n * bh * 4, w4_stride = (size_t)S * bh * 4, h_stride = (size_t)S * 4;
    for (int r = 0; r < brows; ++r) {
        const int a = P.rowptr[(size_t)r], e = P.rowptr[(size_t)r + 1];
        const int sl = r / S, t = r % S;
        const size_t cs = (size_t)s.sbell_slice_ptr[sl], ds = cs * (size_t)bs;
        int last = 0;
        for (int j = a; j < e; ++j) {
            const int c4 = P.bcol[(size_t)j] / 4;
            s.sbell_col_id[cs + t + (size_t)(j - a) * S] = c4;
            last = c4;
            for (int hh = 0; hh < bh; ++hh)
                for (int w4 = 0; w4 < w4n; ++w4)
                    for (int w = 0; w < 4; ++w)
                        s.sbell_data[ds + (size_t)(j - a) * slot_stride + hh * h_stride + w4 * w4_stride + (size_t)t * 4 +
                                     w] = P.data[(size_t)j * bs + hh * bw + w4 * 4 + w];
        }
        for (int j = e; j < a + sn[(size_t)sl]; ++j) s.sbell_col_id[cs + t + (size_t)(j - a) * S] = last;
    }
}

// extract_bell_part (analyze.cpp:1888-2111).
void Analyzer::extract_bell_part(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& rp, int wh, int threshold,
                                 int meth, int bell_num) {
    if (coo.matinfo.nnz == 0) return;
    if (bell_num == 0) {
        copy_coo(coo, rem.m);
        rem.made = true;
        return;
    }
    const int bw = BW[wh], bh = BH[wh], bs = bw * bh;
    const int h = coo.matinfo.height;
    mat.ifusebell = 1;
    mat.bell_meth_num = meth;
    Picked P;
    pick_blocks(coo, rp, bw, bh, [&](int, int nb, int n) { return n > threshold && nb <= bell_num; }, P);
    remainder_from(P, coo.matinfo.width, h, rem);
    auto& b = mat.bell;
    b.matinfo.width = coo.matinfo.width;
    b.matinfo.height = h;
    b.matinfo.nnz = coo.matinfo.nnz - (int)P.rr.size();
    b.b4ell_bwidth = (unsigned)bw;
    b.b4ell_bheight = (unsigned)bh;
    const int brows = (int)P.rowptr.size() - 1;
    b.b4ell_row_num = brows;
    b.b4ell_block_num = bell_num;
    const int Hc = clspmv_aligned_length(brows, kGpuAlign), F = clspmv_aligned_length(4 * brows, kGpuAlign);
    b.b4ell_height_aligned = Hc;
    b.b4ell_float4_aligned = F;
    const int w4n = bw / 4;
    b.b4ell_col_id = xcalloc<int>((size_t)Hc * bell_num);
    b.b4ell_data = xcalloc<float>((size_t)F * bh * w4n * bell_num);
    const size_t slot_stride = (size_t)F * w4n * bh, w4_stride = (size_t)F * bh;
    for (int r = 0; r < brows; ++r) {
        const int a = P.rowptr[(size_t)r], e = P.rowptr[(size_t)r + 1];
        int last = 0;
        for (int j = a; j < e; ++j) {
            const int c4 = P.bcol[(size_t)j] / 4;
            b.b4ell_col_id[r + (size_t)(j - a) * Hc] = c4;
            last = c4;
            for (int hh = 0; hh < bh; ++hh)
                for (int w4 = 0; w4 < w4n; ++w4)
                    for (int w = 0; w < 4; ++w)
                        b.b4ell_data[(size_t)(j - a) * slot_stride + (size_t)hh * F + w4 * w4_stride + (size_t)r * 4 + w] =
                            P.data[(size_t)j * bs + hh * bw + w4 * 4 + w];
        }
        for (int j = e; j < a + bell_num; ++j) b.b4ell_col_id[r + (size_t)(j - a) * Hc] = last;
    }
}

// extract_bcsr_part (analyze.cpp:2113-2296).
void Analyzer::extract_bcsr_part(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& rp, int wh, int bcsr_nnz,
                                 int threshold, int meth) {
    if (coo.matinfo.nnz == 0) return;
    const int bw = BW[wh], bh = BH[wh], bs = bw * bh;
    mat.ifusebcsr = 1;
    mat.bcsr_meth_num = meth;
    Picked P;
    pick_blocks(coo, rp, bw, bh, [&](int, int, int n) { return n > threshold; }, P);
    remainder_from(P, coo.matinfo.width, coo.matinfo.height, rem);
    auto& b = mat.bcsr;
    b.matinfo.width = coo.matinfo.width;
    b.matinfo.height = coo.matinfo.height;
    b.matinfo.nnz = bcsr_nnz;
    b.b4csr_bwidth = (unsigned)bw;
    b.b4csr_bheight = (unsigned)bh;
    const int brows = (int)P.rowptr.size() - 1;
    const size_t nb = P.bcol.size();
    b.b4csr_row_num = brows;
    b.b4csr_row_ptr = xmalloc<int>((size_t)brows + 1);
    std::memcpy(b.b4csr_row_ptr, P.rowptr.data(), ((size_t)brows + 1) * 4);
    b.b4csr_col_id = xmalloc<int>(nb);
    const int N = clspmv_aligned_length((int)nb * 4, kGpuAlign);
    const int w4n = bw / 4;
    b.b4csr_aligned_size = N;
    b.b4csr_block_num = (int)nb;
    b.b4csr_data = xcalloc<float>((size_t)N * bh * w4n);
    const size_t plane = (size_t)N * bh;
    for (size_t i = 0; i < nb; ++i) {
        b.b4csr_col_id[i] = P.bcol[i] / 4;
        for (int hh = 0; hh < bh; ++hh)
            for (int w4 = 0; w4 < w4n; ++w4)
                for (int w = 0; w < 4; ++w)
                    b.b4csr_data[w4 * plane + (size_t)hh * N + i * 4 + w] = P.data[i * bs + hh * bw + w4 * 4 + w];
    }
}

// Slice widths of a block-count vector (per block row) for slices of 32.
static std::vector<int> slice_max(const std::vector<std::pair<int, int>>& bc, int S) {
    const int n = (int)bc.size();
    std::vector<int> out((size_t)(n / S + (n % S != 0)), 0);
    for (size_t i = 0; i < out.size(); ++i) {
        int m = 0;
        for (int j = (int)i * S; j < ((int)i + 1) * S && j < n; ++j) m = std::max(m, bc[(size_t)j].first);
        out[i] = m;
    }
    return out;
}

// extract_full_sbell_bell (analyze.cpp:2435-2519).
void Analyzer::extract_full_sbell_bell(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& rp, int wh,
                                       const std::vector<std::pair<int, int>>& bc, std::pair<int, int> sum) {
    const int h = coo.matinfo.height;
    if (coo.matinfo.nnz == 0) return;
    const int npr = nnz_per_row(sum.second, h);
    maxflop_block(B.sbell, h, npr, info.sbell_meth_num, info.sbell_max_flop);
    maxflop_block(B.bell, h, npr, info.bell_meth_num, info.bell_max_flop);
    int maxb = 0;
    for (auto& p : bc) maxb = std::max(maxb, p.first);
    const int brownum = h / BH[wh] + (h % BH[wh] != 0);
    const double bell_over = overhead_ms(B, brownum / kBellGroup, kBellGroup);
    const int w4n = BW[wh] / 4;
    const int newlength = clspmv_aligned_length(4 * brownum, kGpuAlign);
    double bell_time = 0.0;
    bell_time += ((double)maxb * (double)newlength * (double)BH[wh] * (double)w4n) / 1000000.0 /
                     info.bell_max_flop[wh] + bell_over;
    const int S = kWarp;
    std::vector<int> full = slice_max(bc, S);
    const double sover = overhead_ms(B, brownum / kSellGroup, kSellGroup);
    long long sfull = 0;  // 64-bit, see plan_sell_ell
    for (int v : full) sfull += (long long)S * v;
    const double sbell_full_time = ((double)(sfull * BW[wh] * BH[wh])) / 1000000.0 / info.sbell_max_flop[wh] + sover;
    log("sbell full time %f bell time %f\n", sbell_full_time, bell_time);
    if (bell_time < sbell_full_time) {
        extract_bell_part(coo, rem, rp, wh, 0, info.bell_meth_num[wh], maxb);
        return;
    }
    extract_sbell_part(coo, rem, rp, wh, 0, info.sbell_meth_num[wh], full, S);
}

// extract_sbell_bell (analyze.cpp:2521-2674).
void Analyzer::extract_sbell_bell(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& rp, int wh,
                                  const std::vector<std::pair<int, int>>& bc, std::pair<int, int> sum) {
    const int h = coo.matinfo.height, nnz = coo.matinfo.nnz;
    if (nnz == 0) return;
    const int npr = nnz_per_row(sum.second, h);
    maxflop_block(B.sbell, h, npr, info.sbell_meth_num, info.sbell_max_flop);
    maxflop_block(B.bell, h, npr, info.bell_meth_num, info.bell_max_flop);
    const double rflop = info.max_flat_base_flop;
    const double rover = overhead_ms(B, h / kWorkGroup, kWorkGroup);
    int maxb = 0;
    for (auto& p : bc) maxb = std::max(maxb, p.first);
    std::vector<int> hist((size_t)maxb + 1, 0);
    for (auto& p : bc) hist[(size_t)p.first]++;
    int brownum = h / BH[wh] + (h % BH[wh] != 0);
    int bellnnz = 0;
    double best = 1000000.0;
    int bestnum = 0;
    const double bell_over = overhead_ms(B, brownum / kBellGroup, kBellGroup);
    const int w4n = BW[wh] / 4;
    const int newlength = clspmv_aligned_length(4 * brownum, kGpuAlign);
    const double avgnnz = ((double)sum.second / (double)sum.first);
    for (int i = 0; i < (int)hist.size(); ++i) {
        brownum -= hist[(size_t)i];
        double t = 0.0;
        if (i > 0)
            t += ((double)i * (double)newlength * (double)BH[wh] * (double)w4n) / 1000000.0 / info.bell_max_flop[wh] +
                 bell_over;
        if (bellnnz < nnz) t += ((double)(nnz - bellnnz)) / 1000000.0 / rflop + rover;
        if (t < best) {
            best = t;
            bestnum = i;
        }
        bellnnz += (int)((double)brownum * avgnnz);
    }
    const int S = kWarp;
    brownum = h / BH[wh] + (h % BH[wh] != 0);
    const int ns = brownum / S + (brownum % S != 0);
    std::vector<int> partial((size_t)ns, 0), full((size_t)ns, 0);
    int partial_nnz = 0;
    const int full_nnz = sum.second;
    const double ratio = rflop / info.sbell_max_flop[wh];
    std::vector<int> w;
    for (int i = 0; i < ns; ++i) {
        w.clear();
        for (int j = i * S; j < (i + 1) * S && j < (int)bc.size(); ++j) w.push_back(bc[(size_t)j].first);
        std::sort(w.begin(), w.end());
        full[(size_t)i] = w.back();
        const size_t n = w.size();
        int pn = (int)((double)w[0] * (double)n * avgnnz);
        int cur = pn;
        partial[(size_t)i] = w[0];
        for (size_t j = 1; j < n; ++j) {
            pn += (int)((double)(w[j] - w[j - 1]) * (double)(n - j) * avgnnz);
            if (pn > (int)(((double)(n * (size_t)w[j] * (size_t)BW[wh] * (size_t)BH[wh])) * ratio)) {
                partial[(size_t)i] = w[j];
                cur = pn;
            } else {
                break;
            }
        }
        partial_nnz += cur;
    }
    const double sover = overhead_ms(B, brownum / kSellGroup, kSellGroup);
    long long sfull = 0, spart = 0;  // 64-bit, see plan_sell_ell
    for (int v : full) sfull += (long long)S * v;
    for (int v : partial) spart += (long long)S * v;
    const double full_time = ((double)(sfull * BW[wh] * BH[wh])) / 1000000.0 / info.sbell_max_flop[wh] + sover +
                             ((double)(nnz - full_nnz)) / 1000000.0 / rflop + rover;
    const double part_time = ((double)(spart * BW[wh] * BH[wh])) / 1000000.0 / info.sbell_max_flop[wh] + sover +
                             ((double)(nnz - partial_nnz)) / 1000000.0 / rflop + rover;
    log("sbell full %f partial %f bell %f\n", full_time, part_time, best);
    if (best < full_time && best < part_time) {
        extract_bell_part(coo, rem, rp, wh, 0, info.bell_meth_num[wh], bestnum);
        return;
    }
    if (part_time < full_time) {
        extract_sbell_part(coo, rem, rp, wh, 0, info.sbell_meth_num[wh], partial, S);
        return;
    }
    extract_sbell_part(coo, rem, rp, wh, 0, info.sbell_meth_num[wh], full, S);
}

// extract_block (analyze.cpp:2723-2844) with extract_full_block and
// extract_partial_block (:2676-2721).
bool Analyzer::extract_block(clspmv_coo_matrix& coo, Coo& rem, bool est_partial, bool est_full) {
    const int h = coo.matinfo.height, nnz = coo.matinfo.nnz;
    if (nnz == 0) return false;
    const int npr = nnz_per_row(nnz, h);
    update_block(h, npr);
    update_flat(h, npr);
    const std::vector<int> rp = rowptr_of(coo);
    double best_flop[8];
    int thr[8];
    for (int i = 0; i < 8; ++i) {
        double b = 0.0;
        if (info.sbell_max_flop[i] > b) b = info.sbell_max_flop[i];
        if (info.bell_max_flop[i] > b) b = info.bell_max_flop[i];
        if (info.bcsr_max_flop[i] > b) b = info.bcsr_max_flop[i];
        best_flop[i] = b;
        thr[i] = (int)((double)BH[i] * (double)BW[i] * info.max_flat_base_flop / best_flop[i]);
        if (thr[i] > BH[i] * BW[i] - 1) thr[i] = BH[i] * BW[i] - 1;
    }
    std::vector<std::vector<BlockInfo>> bc;
    count_blocks(coo, rp, thr, bc);
    double block_over[8];
    for (int i = 0; i < 8; ++i) block_over[i] = overhead_ms(B, h / (BH[i] * kSellGroup), kSellGroup);
    const double flatover = overhead_ms(B, h / kWorkGroup, kWorkGroup);
    BlockInfo sum[8];
    for (int i = 0; i < 8; ++i)
        for (const BlockInfo& b : bc[(size_t)i]) {
            sum[i].partial_block_num += b.partial_block_num;
            sum[i].partial_nnz += b.partial_nnz;
            sum[i].full_block_num += b.full_block_num;
            sum[i].full_nnz += b.full_nnz;
        }
    double best_full = 1000000.0, best_part = 1000000.0;
    int full_id = 0, part_id = 0;
    for (int i = 0; i < 8; ++i) {
        const double ft = ((double)sum[i].full_block_num) * ((double)BH[i]) * ((double)BW[i]) / 1000000.0 /
                              best_flop[i] + block_over[i];
        if (ft < best_full) {
            best_full = ft;
            full_id = i;
        }
        std::vector<int> rownnz((size_t)h, 0);
        int flatnnz = 0;
        for (size_t j = 0; j < bc[(size_t)i].size(); ++j) {
            const int diff = bc[(size_t)i][j].full_nnz - bc[(size_t)i][j].partial_nnz;
            const int r0 = (int)j * BH[i];
            const int r1 = std::min(h, ((int)j + 1) * BH[i]);
            const int nr = r1 - r0;
            const int avg = diff / nr;
            for (int k = r0; k < r1 - 1; ++k) rownnz[(size_t)k] = avg;
            rownnz[(size_t)r1 - 1] = diff - avg * (nr - 1);
            flatnnz += diff;
        }
        double pt = ((double)sum[i].partial_block_num) * ((double)BH[i]) * ((double)BW[i]) / 1000000.0 /
                        best_flop[i] + block_over[i];
        if (est_partial)
            pt += estimate_flat_time(rownnz, flatnnz);
        else
            pt += ((double)(nnz - sum[i].partial_nnz)) / 1000000.0 / info.max_flat_base_flop + flatover;
        if (pt < best_part) {
            best_part = pt;
            part_id = i;
        }
        log("block (%d,%d) full %d/%d partial %d/%d full time %f partial time %f\n", BH[i], BW[i],
            sum[i].full_block_num, sum[i].full_nnz, sum[i].partial_block_num, sum[i].partial_nnz, ft, pt);
    }
    double flat_time = ((double)nnz) / 1000000.0 / info.max_flat_base_flop + flatover;
    flat_time *= 1.02;  // favour blocks unless flat is better by 2%
    if (est_full) {
        std::vector<int> rownnz((size_t)h);
        for (int i = 0; i < h; ++i) rownnz[(size_t)i] = rp[(size_t)i + 1] - rp[(size_t)i];
        flat_time = estimate_flat_time(rownnz, nnz);
    }
    if (flat_time <= best_full && flat_time <= best_part) {
        log("No block extracted\n");
        return false;
    }
    if (best_full <= best_part) {
        const int wh = full_id;
        if (info.bcsr_max_flop[wh] >= info.sbell_max_flop[wh] && info.bcsr_max_flop[wh] >= info.bell_max_flop[wh]) {
            extract_bcsr_part(coo, rem, rp, wh, sum[wh].full_nnz, 0, info.bcsr_meth_num[wh]);
            return true;
        }
        std::vector<std::pair<int, int>> cnt(bc[(size_t)wh].size());
        for (size_t i = 0; i < cnt.size(); ++i) cnt[i] = {bc[(size_t)wh][i].full_block_num, bc[(size_t)wh][i].full_nnz};
        extract_full_sbell_bell(coo, rem, rp, wh, cnt, {sum[wh].full_block_num, sum[wh].full_nnz});
        return true;
    }
    const int wh = part_id;
    if (info.bcsr_max_flop[wh] >= info.sbell_max_flop[wh] && info.bcsr_max_flop[wh] >= info.bell_max_flop[wh]) {
        extract_bcsr_part(coo, rem, rp, wh, sum[wh].partial_nnz, thr[wh], info.bcsr_meth_num[wh]);
        return true;
    }
    std::vector<std::pair<int, int>> cnt(bc[(size_t)wh].size());
    for (size_t i = 0; i < cnt.size(); ++i) cnt[i] = {bc[(size_t)wh][i].partial_block_num, bc[(size_t)wh][i].partial_nnz};
    // the reference passes threshold 0 to the partial SBELL/BELL extraction
    extract_sbell_bell(coo, rem, rp, wh, cnt, {sum[wh].partial_block_num, sum[wh].partial_nnz});
    return true;
}

// ---- flat formats (analyze.cpp:2856-3711) ----

void Analyzer::extract_coo(clspmv_coo_matrix& coo) {
    if (coo.matinfo.nnz == 0) return;
    const int npr = nnz_per_row(coo.matinfo.nnz, coo.matinfo.height);
    maxflop_group(B.coo, coo.matinfo.height, npr, info.coo_meth_num, info.coo_groupnum, info.coo_max_flop);
    mat.ifusecoo = 1;
    mat.coo_meth_num = info.coo_meth_num;
    mat.coo_group_num = info.coo_groupnum;
    copy_coo(coo, mat.coo);
}

void Analyzer::extract_csr(clspmv_coo_matrix& coo) {
    if (coo.matinfo.nnz == 0) return;
    const int npr = nnz_per_row(coo.matinfo.nnz, coo.matinfo.height);
    maxflop_group(B.csr, coo.matinfo.height, npr, info.csr_meth_num, info.csr_groupnum, info.csr_max_flop);
    mat.ifusecsr = 1;
    mat.csr_meth_num = info.csr_meth_num;
    mat.csr_group_num = info.csr_groupnum;
    clspmv_coo2csr(&coo, &mat.csr);
}

void Analyzer::extract_csr_coo(clspmv_coo_matrix& coo) {
    const int h = coo.matinfo.height, nnz = coo.matinfo.nnz;
    if (nnz == 0) return;
    const int npr = nnz_per_row(nnz, h);
    maxflop_group(B.csr, h, npr, info.csr_meth_num, info.csr_groupnum, info.csr_max_flop);
    maxflop_group(B.coo, h, npr, info.coo_meth_num, info.coo_groupnum, info.coo_max_flop);
    if (info.coo_max_flop >= info.csr_max_flop) {
        extract_coo(coo);
        return;
    }
    const std::vector<int> cnt = row_counts(coo);
    const double coo_t = ((double)nnz) / 1000000.0 / info.coo_max_flop + overhead_ms(B, info.coo_groupnum, kCooGroup);
    const int newnnz = csr_max_group_nnz(cnt, info.csr_groupnum, info.csr_meth_num) * info.csr_groupnum;
    const double csr_t =
        ((double)newnnz) / 1000000.0 / info.csr_max_flop + overhead_ms(B, info.csr_groupnum, kCsrVecGroup);
    log("coo time %f ms csr time %f ms\n", coo_t, csr_t);
    if (coo_t <= csr_t)
        extract_coo(coo);
    else
        extract_csr(coo);
}

// extract_sell (analyze.cpp:3078-3183): slice widths given; entries past a
// row's slice width go to the remainder (row order, so it stays sorted).
void Analyzer::extract_sell(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& cnt, int S,
                            const std::vector<int>& se, int slice_nnz) {
    const int h = coo.matinfo.height;
    if (coo.matinfo.nnz == 0 || slice_nnz == 0) return;
    mat.ifusesell = 1;
    mat.sell_meth_num = info.sell_meth_num;
    std::vector<int> rp((size_t)h + 1, 0);
    for (int i = 0; i < h; ++i) rp[(size_t)i + 1] = rp[(size_t)i] + cnt[(size_t)i];
    rem.alloc(coo.matinfo.width, h, coo.matinfo.nnz - slice_nnz);
    auto& s = mat.sell;
    const int ns = (int)se.size();
    s.matinfo.width = coo.matinfo.width;
    s.matinfo.height = h;
    s.matinfo.nnz = slice_nnz;
    s.sell_slice_height = (unsigned)S;
    s.sell_slice_num = ns;
    long long slots = 0;
    for (int v : se) slots += (long long)v * S;
    if (slots > INT32_MAX) throw TooLarge("SELL part");
    s.sell_slice_ptr = xmalloc<int>((size_t)ns + 1);
    s.sell_slice_ptr[0] = 0;
    for (int i = 0; i < ns; ++i) s.sell_slice_ptr[i + 1] = s.sell_slice_ptr[i] + se[(size_t)i] * S;
    const size_t total = (size_t)s.sell_slice_ptr[ns];
    s.sell_col_id = xcalloc<int>(total);
    s.sell_data = xcalloc<float>(total);
    int rid = 0;
    for (int i = 0; i < ns; ++i) {
        const int off = i * S;
        int* col = s.sell_col_id + s.sell_slice_ptr[i];
        float* dat = s.sell_data + s.sell_slice_ptr[i];
        const int wdt = se[(size_t)i];
        for (int j = off; j < off + S && j < h; ++j) {
            const int a = rp[(size_t)j], e = rp[(size_t)j + 1];
            for (int k = a; k < e && k < a + wdt; ++k) {
                col[(k - a) * S + j - off] = coo.coo_col_id[k];
                dat[(k - a) * S + j - off] = coo.coo_data[k];
            }
            for (int k = e; k < a + wdt; ++k) {
                const int id = (k - a) * S + j - off;
                col[id] = (id >= S) ? col[id - S] : col[0];
                dat[id] = 0.0f;
            }
            for (int k = a + wdt; k < e; ++k) {
                rem.m.coo_row_id[rid] = j;
                rem.m.coo_col_id[rid] = coo.coo_col_id[k];
                rem.m.coo_data[rid] = coo.coo_data[k];
                ++rid;
            }
        }
        if (i == ns - 1 && ns * S > h)
            for (int j = h; j < ns * S; ++j)
                for (int k = 0; k < wdt; ++k) {
                    const int id = k * S + j - off;
                    col[id] = (id >= S) ? col[id - S] : col[0];
                    dat[id] = 0.0f;
                }
    }
}

// extract_ell (analyze.cpp:3185-3256).
void Analyzer::extract_ell(clspmv_coo_matrix& coo, Coo& rem, const std::vector<int>& cnt, int ellnum, int ellnnz) {
    const int h = coo.matinfo.height;
    if (coo.matinfo.nnz == 0 || ellnum == 0) return;
    maxflop_flat(B.ell, h, ellnum, info.ell_meth_num, info.ell_max_flop);
    mat.ifuseell = 1;
    mat.ell_meth_num = info.ell_meth_num;
    std::vector<int> rp((size_t)h + 1, 0);
    for (int i = 0; i < h; ++i) rp[(size_t)i + 1] = rp[(size_t)i] + cnt[(size_t)i];
    rem.alloc(coo.matinfo.width, h, coo.matinfo.nnz - ellnnz);
    auto& e = mat.ell;
    e.matinfo.width = coo.matinfo.width;
    e.matinfo.height = h;
    e.matinfo.nnz = ellnnz;
    e.ell_num = ellnum;
    const int H = clspmv_aligned_length(h, kGpuAlign);
    e.ell_height_aligned = H;
    e.ell_col_id = xcalloc<int>((size_t)H * ellnum);
    e.ell_data = xcalloc<float>((size_t)H * ellnum);
    int rid = 0;
    for (int i = 0; i < h; ++i) {
        const int a = rp[(size_t)i], en = rp[(size_t)i + 1];
        int last = 0;
        for (int j = a; j < en && j < a + ellnum; ++j) {
            e.ell_col_id[i + (size_t)(j - a) * H] = coo.coo_col_id[j];
            e.ell_data[i + (size_t)(j - a) * H] = coo.coo_data[j];
            last = coo.coo_col_id[j];
        }
        for (int j = en; j < a + ellnum; ++j) e.ell_col_id[i + (size_t)(j - a) * H] = last;
        for (int j = a + ellnum; j < en; ++j) {
            rem.m.coo_row_id[rid] = i;
            rem.m.coo_col_id[rid] = coo.coo_col_id[j];
            rem.m.coo_data[rid] = coo.coo_data[j];
            ++rid;
        }
    }
}

// extract_sell_ell (analyze.cpp:3451-3627).
void Analyzer::extract_sell_ell(clspmv_coo_matrix& coo, Coo& rem) {
    if (coo.matinfo.nnz == 0) return;
    const std::vector<int> cnt = row_counts(coo);
    SellEllPlan p = plan_sell_ell(B, info, cnt, coo.matinfo.nnz, coo.matinfo.height);
    log("ell time %f (width %d) sell full %f partial %f\n", p.ell_time, p.ellnum, p.sell_full_time,
        p.sell_partial_time);
    if (p.ell_time < p.sell_full_time && p.ell_time < p.sell_partial_time) {
        extract_ell(coo, rem, cnt, p.ellnum, p.ellnnz);
        return;
    }
    if (p.sell_partial_time < p.sell_full_time) {
        extract_sell(coo, rem, cnt, p.S, p.slice_partial, p.slice_partial_nnz);
        return;
    }
    extract_sell(coo, rem, cnt, p.S, p.slice_full, p.slice_full_nnz);
}

// extract_flat (analyze.cpp:3675-3711).
void Analyzer::extract_flat(clspmv_coo_matrix& coo) {
    if (coo.matinfo.nnz == 0) return;
    update_flat(coo.matinfo.height, nnz_per_row(coo.matinfo.nnz, coo.matinfo.height));
    if (info.max_flat_base_flop == info.coo_max_flop) {
        extract_coo(coo);
        return;
    }
    if (info.max_flat_base_flop == info.csr_max_flop) {
        extract_csr_coo(coo);
        return;
    }
    Coo rem;
    extract_sell_ell(coo, rem);
    if (!rem.made)
        extract_csr_coo(coo);
    else
        extract_csr_coo(rem.m);
}

// analyze_matrix (analyze.cpp:3714-3900): the three categories are tried in
// the order of their best table GFLOPS; DIA only for square matrices.
void Analyzer::run(clspmv_coo_matrix& coo, float dia_mult, bool est_partial, bool est_full) {
    mat.mat_width = coo.matinfo.width;
    mat.mat_height = coo.matinfo.height;
    mat.mat_nnz = coo.matinfo.nnz;
    const int npr = nnz_per_row(coo.matinfo.nnz, coo.matinfo.height);
    update_dia(coo.matinfo.height, npr);
    update_block(coo.matinfo.height, npr);
    update_flat(coo.matinfo.height, npr);
    double dia_base = info.max_dia_base_flop;
    dia_base *= dia_mult;
    const double block_base = info.max_block_base_flop;
    const double flat_base = info.max_flat_base_flop;
    log("dia flop %f block flop %f flat flop %f\n", dia_base, block_base, flat_base);
    const bool square = coo.matinfo.width == coo.matinfo.height;
    Coo r1, r2;
    auto block_then_flat = [&](clspmv_coo_matrix& in, Coo& rem) {
        if (extract_block(in, rem, est_partial, est_full))
            extract_flat(rem.m);
        else
            extract_flat(in);
    };
    auto dia_then_flat = [&](clspmv_coo_matrix& in, Coo& rem) {
        if (extract_dia(in, rem))
            extract_flat(rem.m);
        else
            extract_flat(in);
    };
    if (dia_base > flat_base) {
        if (block_base > flat_base) {
            if (dia_base > block_base) {
                if (square) {
                    if (extract_dia(coo, r1))
                        block_then_flat(r1.m, r2);
                    else
                        block_then_flat(coo, r2);
                } else {
                    block_then_flat(coo, r1);
                }
            } else {
                if (square) {
                    if (extract_block(coo, r1, est_partial, est_full))
                        dia_then_flat(r1.m, r2);
                    else
                        dia_then_flat(coo, r2);
                } else {
                    block_then_flat(coo, r1);
                }
            }
        } else {
            if (square)
                dia_then_flat(coo, r1);
            else
                extract_flat(coo);
        }
    } else {
        if (block_base > flat_base)
            block_then_flat(coo, r1);
        else
            extract_flat(coo);
    }
}

}  // namespace

extern "C" {

void clspmv_init_cocktail(clspmv_cocktail* mat) { std::memset(mat, 0, sizeof(*mat)); }

void clspmv_free_cocktail(clspmv_cocktail* mat) {
    clspmv_free_bdia(&mat->bdia);
    clspmv_free_dia(&mat->dia);
    clspmv_free_sbell(&mat->sbell);
    clspmv_free_b4ell(&mat->bell);
    clspmv_free_b4csr(&mat->bcsr);
    clspmv_free_sell(&mat->sell);
    clspmv_free_ell(&mat->ell);
    clspmv_free_csr(&mat->csr);
    clspmv_free_coo(&mat->coo);
    mat->ifusebdia = mat->ifusedia = mat->ifusesbell = mat->ifusebell = mat->ifusebcsr = 0;
    mat->ifusesell = mat->ifuseell = mat->ifusecsr = mat->ifusecoo = 0;
}

int clspmv_bench_load(const char* dir, clspmv_bench_t* out) {
    if (!dir || !out) return CLSPMV_E_INVALID;
    auto* B = new clspmv_bench_s();
    const std::string d(dir);
    bool ok = true;
    ok = ok && read_records(d + "/overhead.ben", 2, 1, [&](const double* v) {
        B->over.push_back({(int)v[0], (int)v[1], v[2]});
    });
    ok = ok && read_records(d + "/bdia.ben", 2, kBdiaImp, [&](const double* v) {
        FlatEnt<kBdiaImp> e{(int)v[0], (int)v[1], {}};
        for (int i = 0; i < kBdiaImp; ++i) e.g[i] = v[2 + i];
        B->bdia.push_back(e);
    });
    ok = ok && read_records(d + "/dia.ben", 2, kDiaImp, [&](const double* v) {
        FlatEnt<kDiaImp> e{(int)v[0], (int)v[1], {}};
        for (int i = 0; i < kDiaImp; ++i) e.g[i] = v[2 + i];
        B->dia.push_back(e);
    });
    ok = ok && read_records(d + "/sbell.ben", 4, kSbellImp, [&](const double* v) {
        BlockEnt<kSbellImp> e{(int)v[0], (int)v[1], (int)v[2], (int)v[3], {}};
        for (int i = 0; i < kSbellImp; ++i) e.g[i] = v[4 + i];
        B->sbell.push_back(e);
    });
    ok = ok && read_records(d + "/bell.ben", 4, kBellImp, [&](const double* v) {
        BlockEnt<kBellImp> e{(int)v[0], (int)v[1], (int)v[2], (int)v[3], {}};
        for (int i = 0; i < kBellImp; ++i) e.g[i] = v[4 + i];
        B->bell.push_back(e);
    });
    ok = ok && read_records(d + "/bcsr.ben", 4, kBcsrImp, [&](const double* v) {
        BlockEnt<kBcsrImp> e{(int)v[0], (int)v[1], (int)v[2], (int)v[3], {}};
        for (int i = 0; i < kBcsrImp; ++i) e.g[i] = v[4 + i];
        B->bcsr.push_back(e);
    });
    ok = ok && read_records(d + "/sell.ben", 2, kSellImp, [&](const double* v) {
        FlatEnt<kSellImp> e{(int)v[0], (int)v[1], {}};
        for (int i = 0; i < kSellImp; ++i) e.g[i] = v[2 + i];
        B->sell.push_back(e);
    });
    ok = ok && read_records(d + "/ell.ben", 2, kEllImp, [&](const double* v) {
        FlatEnt<kEllImp> e{(int)v[0], (int)v[1], {}};
        for (int i = 0; i < kEllImp; ++i) e.g[i] = v[2 + i];
        B->ell.push_back(e);
    });
    ok = ok && read_records(d + "/csr.ben", 3, kCsrImp, [&](const double* v) {
        GroupEnt<kCsrImp> e{(int)v[0], (int)v[1], (int)v[2], {}};
        for (int i = 0; i < kCsrImp; ++i) e.g[i] = v[3 + i];
        B->csr.push_back(e);
    });
    ok = ok && read_records(d + "/coo.ben", 3, kCooImp, [&](const double* v) {
        GroupEnt<kCooImp> e{(int)v[0], (int)v[1], (int)v[2], {}};
        for (int i = 0; i < kCooImp; ++i) e.g[i] = v[3 + i];
        B->coo.push_back(e);
    });
    if (!ok || B->over.empty() || B->bdia.empty() || B->dia.empty() || B->sbell.empty() || B->bell.empty() ||
        B->bcsr.empty() || B->sell.empty() || B->ell.empty() || B->csr.empty() || B->coo.empty() ||
        B->bdia.size() != B->dia.size()) {
        delete B;
        set_error("cannot read the performance tables in %s", dir);
        return CLSPMV_E_IO;
    }
    *out = B;
    return CLSPMV_OK;
}

int clspmv_bench_free(clspmv_bench_t b) {
    delete b;
    return CLSPMV_OK;
}

int clspmv_bench_sizes(clspmv_bench_t b, int c[10]) {
    if (!b) return CLSPMV_E_INVALID;
    c[0] = (int)b->over.size();
    c[1] = (int)b->bdia.size();
    c[2] = (int)b->dia.size();
    c[3] = (int)b->sbell.size();
    c[4] = (int)b->bell.size();
    c[5] = (int)b->bcsr.size();
    c[6] = (int)b->sell.size();
    c[7] = (int)b->ell.size();
    c[8] = (int)b->csr.size();
    c[9] = (int)b->coo.size();
    return CLSPMV_OK;
}

int clspmv_maxflop(clspmv_bench_t b, int rows, int nnz, clspmv_maxflop_info* out) {
    if (!b || !out || rows <= 0) return CLSPMV_E_INVALID;
    clspmv_cocktail dummy;
    clspmv_init_cocktail(&dummy);
    Analyzer a(*b, dummy, false);
    const int npr = nnz_per_row(nnz, rows);
    a.update_dia(rows, npr);
    a.update_block(rows, npr);
    a.update_flat(rows, npr);
    *out = a.state();
    return CLSPMV_OK;
}

double clspmv_find_overhead(clspmv_bench_t b, int groupnum, int threadnum) {
    return b ? overhead_us(*b, groupnum, threadnum) : 0.0;
}

int clspmv_analyze_matrix(clspmv_bench_t b, clspmv_coo_matrix* coo, clspmv_cocktail* mat, float dia_mult,
                          int est_partial, int est_full, int verbose) {
    if (!b || !coo || !mat) return CLSPMV_E_INVALID;
    if (coo->matinfo.height <= 0 || coo->matinfo.width <= 0) return CLSPMV_E_INVALID;
    int rc = ensure_sorted(coo);
    if (rc) return rc;
    std::vector<int> rp;
    if ((rc = row_ptr_of(coo, rp))) return rc;
    if ((rc = check_cols(coo))) return rc;
    Analyzer a(*b, *mat, verbose != 0);
    try {
        a.run(*coo, dia_mult, est_partial != 0, est_full != 0);
    } catch (const TooLarge& e) {
        clspmv_free_cocktail(mat);
        set_error("analyze_matrix: %s exceeds 2^31 stored entries", e.what);
        return CLSPMV_E_TOO_LARGE;
    }
    return CLSPMV_OK;
}

}  // extern "C"
