// Internal definitions shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

struct cav_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    void* render = nullptr;  // render-path state owned by the context (render.cpp)
};

namespace cav {
int fail(cav_ctx* ctx, int code, const std::string& msg);
int hip_fail(cav_ctx* ctx, hipError_t e, const char* what);
int dev_alloc(cav_ctx* ctx, void** p, size_t bytes, const char* what);
// frees context-owned state of other subsystems before the stream goes away
void ctx_release(cav_ctx* ctx);
}  // namespace cav
