// Kernel launches that either run on a stream or are recorded as nodes of
// a HIP graph under construction.
//
// The multigrid plans replay fixed launch sequences as graphs.  Building
// those graphs by stream capture is fragile in a multi-threaded host: on
// HIP, a thread whose capture mode is the default (global) that makes an
// allocation or a synchronising call while any stream of the process is
// capturing invalidates that capture, and torch or other serving threads
// make such calls at any time.  Here graphs are assembled explicitly
// instead: while a GraphRecorder is active on the calling thread, every
// launch_kernel() call appends a kernel node that depends on the previous
// node (the stream order of the recorded sequence); nothing is captured, so
// nothing can be invalidated by another thread.
#pragma once
#include <hip/hip_runtime.h>

#include <tuple>
#include <type_traits>
#include <utility>

namespace gpaw {

struct GraphRecorder {
    hipGraph_t graph = nullptr;
    hipGraphNode_t last = nullptr;
    hipError_t err = hipSuccess;
};

// the recorder of the calling thread (nullptr: launch directly)
GraphRecorder*& graph_recorder();

namespace detail {
template <class Tuple, size_t... I>
void arg_ptrs(Tuple& t, void** p, std::index_sequence<I...>)
{
    ((p[I] = (void*)&std::get<I>(t)), ...);
}
}  // namespace detail

template <class... P, class... A>
hipError_t launch_kernel(void (*k)(P...), dim3 grid, dim3 block, size_t shmem, hipStream_t st,
                         A&&... a)
{
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    std::tuple<std::decay_t<P>...> vals(std::forward<A>(a)...);
    void* ptrs[sizeof...(P) + 1];
    detail::arg_ptrs(vals, ptrs, std::index_sequence_for<P...>{});
    GraphRecorder* rec = graph_recorder();
    if (!rec) return hipLaunchKernel((const void*)k, grid, block, ptrs, shmem, st);
    if (rec->err != hipSuccess) return rec->err;
    hipKernelNodeParams kp{};
    kp.func = (void*)k;
    kp.gridDim = grid;
    kp.blockDim = block;
    kp.sharedMemBytes = (unsigned)shmem;
    kp.kernelParams = ptrs;
    kp.extra = nullptr;
    hipGraphNode_t node;
    rec->err = hipGraphAddKernelNode(&node, rec->graph, rec->last ? &rec->last : nullptr,
                                     rec->last ? 1 : 0, &kp);
    if (rec->err == hipSuccess) rec->last = node;
    return rec->err;
}

// Status of the launches since the last call: the recorder's error while
// recording, hipGetLastError() otherwise.
inline hipError_t launch_error()
{
    GraphRecorder* rec = graph_recorder();
    return rec ? rec->err : hipGetLastError();
}

// Record body(stream) into a new executable graph.  body must only issue
// launch_kernel() calls (no copies, memsets or synchronisation).
template <class F>
hipError_t build_graph(hipGraphExec_t* out, F&& body)
{
    *out = nullptr;
    GraphRecorder rec;
    hipError_t e = hipGraphCreate(&rec.graph, 0);
    if (e != hipSuccess) return e;
    GraphRecorder* prev = graph_recorder();
    graph_recorder() = &rec;
    hipError_t e1 = body((hipStream_t) nullptr);
    graph_recorder() = prev;
    if (e1 == hipSuccess) e1 = rec.err;
    if (e1 == hipSuccess) e1 = hipGraphInstantiate(out, rec.graph, nullptr, nullptr, 0);
    hipGraphDestroy(rec.graph);
    if (e1 != hipSuccess) *out = nullptr;
    return e1;
}

}  // namespace gpaw
