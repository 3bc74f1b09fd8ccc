// Closest-hit traversal that allocates nothing.
//
// cav_trace_closest (host buffers) allocates device staging for the rays and
// hits on every call. Callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc)
// use cav_trace_closest_noalloc instead: the same kernels, staged through one
// file-scope device array per GPU. The array's device address is looked up
// once per device, with the device current, when a context is created on it
// (noalloc_init, called by cav_ctx_create); host code only ever passes that
// address, never the array's name, to kernels and copies.
//
// Layout of the array: 256 bytes of traversal counters (overflow, ray
// dispensers, deepest stack, fetched records), then the rays (24 B each),
// hit prim, t, u and v (4 B each) of up to CAV_NOALLOC_MAX_RAYS rays.
// Every call on one GPU shares that GPU's array and leaves the previous
// call's contents in it; each call clears the counters before its launch.
// Calls that overlap on one GPU (from several contexts or threads) race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "capi_internal.h"
#include "lbvh.h"

namespace {

constexpr int64_t kMaxRays = CAV_NOALLOC_MAX_RAYS;
constexpr size_t kCounterBytes = 256;
constexpr size_t kScratchBytes = kCounterBytes + (size_t)kMaxRays * (24 + 4 + 4 + 4 + 4);

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::once_flag g_once[kMaxDevices];
void* g_ptr[kMaxDevices];           // device address of `scratch` on each device
hipError_t g_err[kMaxDevices];      // result of its lookup

}  // namespace

namespace cav {

// Looks up the scratch array's address on `device`, which must be current.
// Once per device; later calls return the first lookup's result.
hipError_t noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::call_once(g_once[device], [device] {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_err[device] = e;
        g_ptr[device] = e == hipSuccess ? p : nullptr;
    });
    return g_err[device];
}

}  // namespace cav

#define NACHK(ctx, call)                                               \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return cav::hip_fail((ctx), e_, #call);  \
    } while (0)

extern "C" int cav_trace_closest_noalloc(cav_ctx* ctx, const cav_bvh* h, const float* rays, int64_t nrays,
                                         int32_t* prim, float* t, float* u, float* v) {
    if (!ctx || !h) return CAV_ERR_INVALID;
    if (nrays < 0 || (nrays > 0 && (!rays || !prim || !t)))
        return cav::fail(ctx, CAV_ERR_INVALID, "null ray or output buffer");
    if (nrays > kMaxRays)
        return cav::fail(ctx, CAV_ERR_INVALID,
                         "cav_trace_closest_noalloc takes at most " + std::to_string(kMaxRays) +
                             " rays per call (got " + std::to_string(nrays) + ")");
    if (nrays == 0) return CAV_OK;
    const int dev = ctx->device;
    if (dev < 0 || dev >= kMaxDevices || !g_ptr[dev])
        return cav::fail(ctx, CAV_ERR_HIP,
                         std::string("no-alloc scratch unavailable on this device: ") +
                             (dev >= 0 && dev < kMaxDevices ? hipGetErrorString(g_err[dev]) : "bad device"));
    NACHK(ctx, hipSetDevice(dev));
    hipStream_t s = ctx->stream;
    unsigned char* base = static_cast<unsigned char*>(g_ptr[dev]);
    uint32_t* counters = reinterpret_cast<uint32_t*>(base);
    float* drays = reinterpret_cast<float*>(base + kCounterBytes);
    int32_t* dprim = reinterpret_cast<int32_t*>(drays + 6 * kMaxRays);
    float* dt = reinterpret_cast<float*>(dprim + kMaxRays);
    float* du = dt + kMaxRays;
    float* dv = du + kMaxRays;
    const size_t r = (size_t)nrays;
    NACHK(ctx, hipMemcpyAsync(drays, rays, 24 * r, hipMemcpyHostToDevice, s));
    NACHK(ctx, hipMemsetAsync(counters, 0, kCounterBytes, s));
    NACHK(ctx, cav::trace_closest(h->rec, h->n, h->root, drays, nrays, dprim, dt, u ? du : nullptr,
                                  v ? dv : nullptr, counters, ctx->launch[CAV_TRACE_STANDALONE], s));
    uint32_t ovf = 0;
    NACHK(ctx, hipMemcpyAsync(&ovf, counters, 4, hipMemcpyDeviceToHost, s));
    NACHK(ctx, hipMemcpyAsync(prim, dprim, 4 * r, hipMemcpyDeviceToHost, s));
    NACHK(ctx, hipMemcpyAsync(t, dt, 4 * r, hipMemcpyDeviceToHost, s));
    if (u) NACHK(ctx, hipMemcpyAsync(u, du, 4 * r, hipMemcpyDeviceToHost, s));
    if (v) NACHK(ctx, hipMemcpyAsync(v, dv, 4 * r, hipMemcpyDeviceToHost, s));
    NACHK(ctx, hipStreamSynchronize(s));
    if (ovf) return cav::fail(ctx, CAV_ERR_OVERFLOW, "traversal stack overflow in " + std::to_string(ovf) + " pushes");
    return CAV_OK;
}
