// Internal runtime types of libclspmv (context, device buffers).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "clspmv.h"

struct clspmv_ctx_s {
    int device = 0;
    hipStream_t own = nullptr;  // default-priority stream created by clspmv_init
    hipStream_t ext = nullptr;  // caller's stream, if set
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipStream_t stream() const { return ext ? ext : own; }
};

namespace clspmv {

int hip_status(hipError_t e, const char* what);

// One device allocation owned by a part or a Cocktail object.
struct DevBuf {
    void* ptr = nullptr;
    size_t size = 0;
    int upload(clspmv_ctx_t ctx, const void* host, size_t bytes);
    int alloc(size_t bytes);
    void release();
};

}  // namespace clspmv

extern "C" int clspmv_part_launch(clspmv_part_t p, const float* x, float* y, hipStream_t st);
