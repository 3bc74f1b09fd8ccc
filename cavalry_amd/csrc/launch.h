// Kernel launches that go either to a stream or into a HIP graph.
//
// A graph is built node by node (hipGraphAddKernelNode), each node depending
// on the one added before it, so replaying it runs the kernels in the order
// they were "launched". Nothing is captured from a stream: graph building
// involves no stream state, so HIP calls made by other threads at the same
// time (allocations, synchronous copies) cannot invalidate or fail it.
#pragma once
#include <hip/hip_runtime.h>

#include <tuple>
#include <utility>

namespace cav {

struct LaunchSink {
    hipStream_t stream = nullptr;  // direct launches
    hipGraph_t graph = nullptr;    // or nodes appended to this graph
    hipGraphNode_t last = nullptr;
    hipError_t err = hipSuccess;
    const char* failed = nullptr;  // name of the first launch that failed

    explicit LaunchSink(hipStream_t s) : stream(s) {}
    explicit LaunchSink(hipGraph_t g) : graph(g) {}

    // Launches kernel(args...) on grid x block threads (no dynamic LDS). The
    // arguments are converted to the kernel's parameter types first; a graph
    // node copies them, so they need not outlive the call.
    template <typename... P, typename... A>
    void launch(const char* name, void (*kernel)(P...), dim3 grid, dim3 block, A&&... args) {
        static_assert(sizeof...(P) == sizeof...(A), "argument count");
        if (err != hipSuccess) return;
        std::tuple<P...> vals(std::forward<A>(args)...);
        go(name, reinterpret_cast<const void*>(kernel), grid, block, vals, std::index_sequence_for<P...>{});
    }

private:
    template <typename T, size_t... I>
    void go(const char* name, const void* fn, dim3 grid, dim3 block, T& vals, std::index_sequence<I...>) {
        void* params[sizeof...(I) > 0 ? sizeof...(I) : 1] = {static_cast<void*>(&std::get<I>(vals))...};
        hipError_t e;
        if (graph) {
            hipKernelNodeParams np{};
            np.func = const_cast<void*>(fn);
            np.gridDim = grid;
            np.blockDim = block;
            np.sharedMemBytes = 0;
            np.kernelParams = params;
            np.extra = nullptr;
            hipGraphNode_t node = nullptr;
            e = hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &np);
            if (e == hipSuccess) last = node;
        } else {
            e = hipLaunchKernel(fn, grid, block, params, 0, stream);
        }
        if (e != hipSuccess) {
            err = e;
            failed = name;
        }
    }
};

}  // namespace cav
