// The selector network: forward, softmax cross-entropy, backward and Adam,
// as hand-written HIP kernels (FP32).  Sizes are small (tens of features,
// 128 hidden units, a few thousand samples), so one LDS-tiled GEMM kernel
// with transpose flags serves every product, and the reductions (bias
// gradients, the loss) are column loops in a fixed order: training is
// deterministic for a given seed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <new>
#include <vector>

#include "clspmv.h"
#include "clspmv_mlp.h"
#include "host/internal.hpp"
#include "runtime.hpp"

using namespace clspmv;

namespace {

constexpr int kT = 16;  // GEMM tile

// C[M x N] = op(A) * op(B) (+ C when accumulate), op(A) is M x K, op(B) K x N.
// A is row-major M x K (lda = K) or, transposed, K x M (lda = M); likewise B.
__global__ __launch_bounds__(kT* kT) void gemm_kernel(int M, int N, int K, const float* __restrict__ A, int ta,
                                                      const float* __restrict__ B, int tb, float* __restrict__ C,
                                                      int accumulate) {
    __shared__ float As[kT][kT + 1];
    __shared__ float Bs[kT][kT + 1];
    const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;
    const int row = blockIdx.y * kT + ty, col = blockIdx.x * kT + tx;
    float acc = 0.f;
    for (int k0 = 0; k0 < K; k0 += kT) {
        const int ka = k0 + tx, kb = k0 + ty;
        float a = 0.f, b = 0.f;
        if (row < M && ka < K) a = ta ? A[(size_t)ka * M + row] : A[(size_t)row * K + ka];
        if (col < N && kb < K) b = tb ? B[(size_t)col * K + kb] : B[(size_t)kb * N + col];
        As[ty][tx] = a;
        Bs[ty][tx] = b;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kT; ++k) acc += As[ty][k] * Bs[k][tx];
        __syncthreads();
    }
    if (row < M && col < N) {
        float* c = C + (size_t)row * N + col;
        *c = accumulate ? *c + acc : acc;
    }
}

__global__ void bias_act_kernel(float* __restrict__ Z, const float* __restrict__ b, int M, int N, int relu) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= M * N) return;
    float v = Z[i] + b[i % N];
    Z[i] = relu ? fmaxf(v, 0.f) : v;
}

// Per sample: softmax of the logits; dZ = (p - onehot) / n; loss_i.
__global__ void softmax_ce_kernel(const float* __restrict__ Z, const int* __restrict__ y, int n, int c,
                                  float* __restrict__ dZ, float* __restrict__ loss_i) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float* z = Z + (size_t)i * c;
    float mx = z[0];
    for (int j = 1; j < c; ++j) mx = fmaxf(mx, z[j]);
    float s = 0.f;
    for (int j = 0; j < c; ++j) s += expf(z[j] - mx);
    const int t = y[i];
    for (int j = 0; j < c; ++j) {
        const float p = expf(z[j] - mx) / s;
        dZ[(size_t)i * c + j] = (p - (j == t ? 1.f : 0.f)) / (float)n;
    }
    loss_i[i] = -(z[t] - mx - logf(s));
}

__global__ void relu_back_kernel(float* __restrict__ dH, const float* __restrict__ H, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n && H[i] <= 0.f) dH[i] = 0.f;
}

// out[j] = sum_i A[i][j], i ascending (fixed order).
__global__ void colsum_kernel(const float* __restrict__ A, int M, int N, float* __restrict__ out) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= N) return;
    float s = 0.f;
    for (int i = 0; i < M; ++i) s += A[(size_t)i * N + j];
    out[j] = s;
}

__global__ void sum_kernel(const float* __restrict__ a, int n, float* __restrict__ out) {
    if (threadIdx.x || blockIdx.x) return;
    float s = 0.f;
    for (int i = 0; i < n; ++i) s += a[i];
    *out = s / (float)n;
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long long n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void argmax_kernel(const float* __restrict__ Z, int n, int c, int* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int best = 0;
    for (int j = 1; j < c; ++j)
        if (Z[(size_t)i * c + j] > Z[(size_t)i * c + best]) best = j;
    out[i] = best;
}

dim3 blocks(long long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

struct clspmv_mlp_s {
    clspmv_ctx_t ctx = nullptr;
    int in = 0, hid = 0, cls = 0;
    long long np = 0;
    DevBuf p, g, m, v;          // parameters, gradients, Adam moments
    long long t = 0;            // Adam steps taken since creation or the last set_params
    DevBuf x, y, h1, h2, z;     // activations for the current n
    DevBuf d1, d2, dz, loss_i, loss;
    int cap = 0;                // samples the activation buffers hold
    float *W1 = nullptr, *b1 = nullptr, *W2 = nullptr, *b2 = nullptr, *W3 = nullptr, *b3 = nullptr;
    float *gW1 = nullptr, *gb1 = nullptr, *gW2 = nullptr, *gb2 = nullptr, *gW3 = nullptr, *gb3 = nullptr;
    ~clspmv_mlp_s() {
        for (DevBuf* b : {&p, &g, &m, &v, &x, &y, &h1, &h2, &z, &d1, &d2, &dz, &loss_i, &loss}) b->release();
    }
    void slice(float* base, float*& w1, float*& c1, float*& w2, float*& c2, float*& w3, float*& c3) const {
        w1 = base;
        c1 = w1 + (size_t)in * hid;
        w2 = c1 + hid;
        c2 = w2 + (size_t)hid * hid;
        w3 = c2 + hid;
        c3 = w3 + (size_t)hid * cls;
    }
    hipStream_t st() const { return ctx->stream(); }
    int ensure(int n) {
        if (n <= cap) return CLSPMV_OK;
        for (DevBuf* b : {&x, &y, &h1, &h2, &z, &d1, &d2, &dz, &loss_i}) b->release();
        int rc = x.alloc((size_t)n * in * 4);
        if (!rc) rc = y.alloc((size_t)n * 4);
        if (!rc) rc = h1.alloc((size_t)n * hid * 4);
        if (!rc) rc = h2.alloc((size_t)n * hid * 4);
        if (!rc) rc = z.alloc((size_t)n * cls * 4);
        if (!rc) rc = d1.alloc((size_t)n * hid * 4);
        if (!rc) rc = d2.alloc((size_t)n * hid * 4);
        if (!rc) rc = dz.alloc((size_t)n * cls * 4);
        if (!rc) rc = loss_i.alloc((size_t)n * 4);
        if (!rc) cap = n;
        return rc;
    }
    void gemm(int M, int N, int K, const float* A, int ta, const float* B, int tb, float* C, int acc) {
        hipLaunchKernelGGL(gemm_kernel, dim3((N + kT - 1) / kT, (M + kT - 1) / kT), dim3(kT * kT), 0, st(), M, N,
                           K, A, ta, B, tb, C, acc);
    }
    void forward(int n) {
        const float* X = (const float*)x.ptr;
        float *H1 = (float*)h1.ptr, *H2 = (float*)h2.ptr, *Z = (float*)z.ptr;
        gemm(n, hid, in, X, 0, W1, 0, H1, 0);
        hipLaunchKernelGGL(bias_act_kernel, blocks((long long)n * hid), dim3(256), 0, st(), H1, b1, n, hid, 1);
        gemm(n, hid, hid, H1, 0, W2, 0, H2, 0);
        hipLaunchKernelGGL(bias_act_kernel, blocks((long long)n * hid), dim3(256), 0, st(), H2, b2, n, hid, 1);
        gemm(n, cls, hid, H2, 0, W3, 0, Z, 0);
        hipLaunchKernelGGL(bias_act_kernel, blocks((long long)n * cls), dim3(256), 0, st(), Z, b3, n, cls, 0);
    }
    void backward(int n) {
        const float* X = (const float*)x.ptr;
        float *H1 = (float*)h1.ptr, *H2 = (float*)h2.ptr, *Z = (float*)z.ptr;
        float *D1 = (float*)d1.ptr, *D2 = (float*)d2.ptr, *DZ = (float*)dz.ptr;
        hipLaunchKernelGGL(softmax_ce_kernel, blocks(n), dim3(256), 0, st(), Z, (const int*)y.ptr, n, cls, DZ,
                           (float*)loss_i.ptr);
        hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, st(), (const float*)loss_i.ptr, n, (float*)loss.ptr);
        gemm(hid, cls, n, H2, 1, DZ, 0, gW3, 0);                       // dW3 = H2^T dZ
        hipLaunchKernelGGL(colsum_kernel, blocks(cls), dim3(256), 0, st(), DZ, n, cls, gb3);
        gemm(n, hid, cls, DZ, 0, W3, 1, D2, 0);                        // dH2 = dZ W3^T
        hipLaunchKernelGGL(relu_back_kernel, blocks((long long)n * hid), dim3(256), 0, st(), D2, H2, n * hid);
        gemm(hid, hid, n, H1, 1, D2, 0, gW2, 0);                       // dW2 = H1^T dH2
        hipLaunchKernelGGL(colsum_kernel, blocks(hid), dim3(256), 0, st(), D2, n, hid, gb2);
        gemm(n, hid, hid, D2, 0, W2, 1, D1, 0);                        // dH1 = dH2 W2^T
        hipLaunchKernelGGL(relu_back_kernel, blocks((long long)n * hid), dim3(256), 0, st(), D1, H1, n * hid);
        gemm(in, hid, n, X, 1, D1, 0, gW1, 0);                         // dW1 = X^T dH1
        hipLaunchKernelGGL(colsum_kernel, blocks(hid), dim3(256), 0, st(), D1, n, hid, gb1);
    }
    int upload(const float* hx, const int* hy, int n) {
        int rc = ensure(n);
        if (!rc) rc = hip_status(hipMemcpyAsync(x.ptr, hx, (size_t)n * in * 4, hipMemcpyHostToDevice, st()), "h2d");
        if (!rc && hy) rc = hip_status(hipMemcpyAsync(y.ptr, hy, (size_t)n * 4, hipMemcpyHostToDevice, st()), "h2d");
        return rc;
    }
};

extern "C" {

int clspmv_mlp_create(clspmv_ctx_t ctx, int in, int hidden, int classes, uint64_t seed, clspmv_mlp_t* out) {
    if (!ctx || !out || in <= 0 || hidden <= 0 || classes <= 1 || hidden > 4096 || in > 4096) return CLSPMV_E_INVALID;
    auto* m = new (std::nothrow) clspmv_mlp_s();
    if (!m) return CLSPMV_E_NOMEM;
    m->ctx = ctx;
    m->in = in;
    m->hid = hidden;
    m->cls = classes;
    m->np = (long long)in * hidden + hidden + (long long)hidden * hidden + hidden + (long long)hidden * classes + classes;
    // He-uniform weights from a SplitMix64 stream, zero biases
    std::vector<float> h((size_t)m->np, 0.f);
    uint64_t s = seed;
    auto next = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](size_t off, long long fan_in, long long count) {
        const double a = std::sqrt(6.0 / (double)fan_in);
        for (long long i = 0; i < count; ++i) h[off + (size_t)i] = (float)((2.0 * next() - 1.0) * a);
    };
    size_t off = 0;
    fill(off, in, (long long)in * hidden);
    off += (size_t)in * hidden + hidden;
    fill(off, hidden, (long long)hidden * hidden);
    off += (size_t)hidden * hidden + hidden;
    fill(off, hidden, (long long)hidden * classes);
    int rc = m->p.upload(ctx, h.data(), h.size() * 4);
    std::fill(h.begin(), h.end(), 0.f);
    if (!rc) rc = m->g.upload(ctx, h.data(), h.size() * 4);
    if (!rc) rc = m->m.upload(ctx, h.data(), h.size() * 4);
    if (!rc) rc = m->v.upload(ctx, h.data(), h.size() * 4);
    if (!rc) rc = m->loss.alloc(4);
    if (rc) {
        delete m;
        return rc;
    }
    m->slice((float*)m->p.ptr, m->W1, m->b1, m->W2, m->b2, m->W3, m->b3);
    m->slice((float*)m->g.ptr, m->gW1, m->gb1, m->gW2, m->gb2, m->gW3, m->gb3);
    *out = m;
    return CLSPMV_OK;
}

int clspmv_mlp_free(clspmv_mlp_t m) {
    if (!m) return CLSPMV_E_INVALID;
    (void)hipStreamSynchronize(m->st());
    delete m;
    return CLSPMV_OK;
}

long long clspmv_mlp_num_params(clspmv_mlp_t m) { return m ? m->np : 0; }

int clspmv_mlp_get_params(clspmv_mlp_t m, float* params) {
    if (!m || !params) return CLSPMV_E_INVALID;
    return clspmv_memcpy_d2h(m->ctx, params, m->p.ptr, (size_t)m->np * 4);
}
// New parameters start a new optimisation: the Adam moments and step count
// are reset with them.
int clspmv_mlp_set_params(clspmv_mlp_t m, const float* params) {
    if (!m || !params) return CLSPMV_E_INVALID;
    int rc = clspmv_memcpy_h2d(m->ctx, m->p.ptr, params, (size_t)m->np * 4);
    if (!rc) rc = hip_status(hipMemsetAsync(m->m.ptr, 0, (size_t)m->np * 4, m->st()), "hipMemset");
    if (!rc) rc = hip_status(hipMemsetAsync(m->v.ptr, 0, (size_t)m->np * 4, m->st()), "hipMemset");
    if (!rc) m->t = 0;
    return rc;
}

static int check_labels(const clspmv_mlp_s* m, const int* labels, int n) {
    for (int i = 0; i < n; ++i)
        if (labels[i] < 0 || labels[i] >= m->cls) {
            set_error("label %d of sample %d outside [0, %d)", labels[i], i, m->cls);
            return CLSPMV_E_INVALID;
        }
    return CLSPMV_OK;
}

int clspmv_mlp_loss_grad(clspmv_mlp_t m, const float* x, const int* labels, int n, float* loss, float* grads) {
    if (!m || !x || !labels || n <= 0) return CLSPMV_E_INVALID;
    int rc = check_labels(m, labels, n);
    if (!rc) rc = m->upload(x, labels, n);
    if (rc) return rc;
    m->forward(n);
    m->backward(n);
    rc = hip_status(hipGetLastError(), "mlp kernels");
    if (!rc && loss) rc = clspmv_memcpy_d2h(m->ctx, loss, m->loss.ptr, 4);
    if (!rc && grads) rc = clspmv_memcpy_d2h(m->ctx, grads, m->g.ptr, (size_t)m->np * 4);
    return rc;
}

int clspmv_mlp_train(clspmv_mlp_t m, const float* x, const int* labels, int n, int steps, float lr, float beta1,
                     float beta2, float eps, float* loss) {
    if (!m || !x || !labels || n <= 0 || steps < 0) return CLSPMV_E_INVALID;
    int rc = check_labels(m, labels, n);
    if (!rc) rc = m->upload(x, labels, n);
    if (rc) return rc;
    // the step count continues across calls, as the moments do, so that a
    // run split over several calls takes the same steps as one call
    for (int s = 0; s < steps; ++s) {
        m->forward(n);
        m->backward(n);
        const long long t = ++m->t;
        const float c1 = (float)(1.0 - std::pow((double)beta1, (double)t));
        const float c2 = (float)(1.0 - std::pow((double)beta2, (double)t));
        hipLaunchKernelGGL(adam_kernel, blocks(m->np), dim3(256), 0, m->st(), (float*)m->p.ptr,
                           (const float*)m->g.ptr, (float*)m->m.ptr, (float*)m->v.ptr, m->np, lr, beta1, beta2, eps,
                           c1, c2);
    }
    rc = hip_status(hipGetLastError(), "mlp train");
    if (!rc) {
        m->forward(n);
        m->backward(n);
    }
    if (!rc && loss) rc = clspmv_memcpy_d2h(m->ctx, loss, m->loss.ptr, 4);
    return rc;
}

int clspmv_mlp_predict(clspmv_mlp_t m, const float* x, int n, float* logits, int* classes) {
    if (!m || !x || n <= 0) return CLSPMV_E_INVALID;
    int rc = m->upload(x, nullptr, n);
    if (rc) return rc;
    m->forward(n);
    hipLaunchKernelGGL(argmax_kernel, blocks(n), dim3(256), 0, m->st(), (const float*)m->z.ptr, n, m->cls,
                       (int*)m->y.ptr);
    rc = hip_status(hipGetLastError(), "mlp predict");
    if (!rc && logits) rc = clspmv_memcpy_d2h(m->ctx, logits, m->z.ptr, (size_t)n * m->cls * 4);
    if (!rc && classes) rc = clspmv_memcpy_d2h(m->ctx, classes, m->y.ptr, (size_t)n * 4);
    return rc;
}

}  // extern "C"
