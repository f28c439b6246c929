// Kernel launches that can also be recorded as the nodes of an explicitly
// built HIP graph.  Blocks of Krylov iterations are turned into graphs this
// way rather than by stream capture: a capture in progress on one host
// thread makes the runtime reject legacy-stream work issued by any other
// thread (a synchronous hipMemcpy, for one), which a server running solves
// on several threads cannot rule out.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <tuple>
#include <type_traits>
#include <utility>

namespace ehip {

// While a recorder is installed on this thread, launch_k adds a kernel node
// (depending on the previous one) to its graph instead of launching.
struct GraphRec {
  hipGraph_t graph = nullptr;
  hipGraphNode_t last = nullptr;
  hipError_t err = hipSuccess;
};
extern thread_local GraphRec *tl_graph_rec;

// First failure of a direct launch_k on this thread since the last
// take_launch_error() (HIP's own last-error slot is overwritten by every
// later call, so launch failures are kept here until a driver checks them).
extern thread_local hipError_t tl_launch_err;
inline hipError_t take_launch_error() {
  const hipError_t e = tl_launch_err;
  tl_launch_err = hipSuccess;
  return e;
}

// Whether iteration blocks may be replayed as graphs: off when EHIP_GRAPHS=0,
// and off under rocprofv3 (ROCPROF_OUTPUT_PATH in the environment), where
// the same kernels are launched one by one instead.
bool graphs_enabled();

template <typename Tuple, std::size_t... I>
inline void kernel_arg_ptrs(Tuple &t, void **out, std::index_sequence<I...>) {
  ((out[I] = (void *)&std::get<I>(t)), ...);
}

template <typename... KArgs, typename... Args>
inline void launch_k(void (*kernel)(KArgs...), dim3 grid, dim3 block, hipStream_t stream, Args &&...args) {
  static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
  std::tuple<std::decay_t<KArgs>...> a(std::forward<Args>(args)...);
  void *ptrs[sizeof...(KArgs)];
  kernel_arg_ptrs(a, ptrs, std::index_sequence_for<KArgs...>{});
  GraphRec *r = tl_graph_rec;
  if (!r) {
    const hipError_t e = hipLaunchKernel(reinterpret_cast<const void *>(kernel), grid, block, ptrs, 0, stream);
    if (e != hipSuccess && tl_launch_err == hipSuccess) tl_launch_err = e;
    return;
  }
  if (r->err != hipSuccess) return;
  hipKernelNodeParams p{};
  p.func = reinterpret_cast<void *>(kernel);
  p.gridDim = grid;
  p.blockDim = block;
  p.sharedMemBytes = 0;
  p.kernelParams = ptrs;  // the node keeps its own copy of the argument values
  p.extra = nullptr;
  hipGraphNode_t node = nullptr;
  r->err = hipGraphAddKernelNode(&node, r->graph, r->last ? &r->last : nullptr, r->last ? 1 : 0, &p);
  if (r->err == hipSuccess) r->last = node;
}

}  // namespace ehip
