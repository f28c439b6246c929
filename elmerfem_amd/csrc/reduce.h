// Deterministic wave64 / workgroup reductions and the "last workgroup
// finishes" pattern used by every fused Krylov kernel.  No floating-point
// atomics: workgroups publish partial sums, take an integer ticket, and the
// workgroup holding the final ticket sums all partials in index order.
#pragma once
#include <hip/hip_runtime.h>

namespace ehip {

constexpr int kWave = 64;

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  // butterfly over the 64 lanes; the pairing is fixed, so the rounding is too
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sums one value per thread over the workgroup (blockDim.x <= 1024).
// Every thread receives the total.  `sh` needs blockDim.x/64 elements.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int nw = (blockDim.x + kWave - 1) / kWave;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T t = (T)0;
  for (int i = 0; i < nw; ++i) t += sh[i];
  return t;
}

template <typename T>
__device__ __forceinline__ T load_coherent(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Writes the workgroup's NV partial sums to partials[k*stride + blockIdx.x]
// with plain stores.  A later launch on the same stream (the one-workgroup
// finalize kernel) reads them: kernel boundaries order the memory, so no
// fence is needed here.  (An agent-scope fence on gfx950 writes back and
// invalidates the whole L2 of the XCD -- per workgroup that costs more than
// the reduction saves.)
template <typename T, int NV>
__device__ __forceinline__ void block_publish(const T (&v)[NV], T *partials, int stride) {
  __shared__ T sh[16 * NV];
  T bs[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) bs[k] = block_sum(v[k], sh + 16 * k);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * stride + blockIdx.x] = bs[k];
  }
}

// Sums g published partials per value in a fixed order; one workgroup, every
// thread receives the totals.
template <typename T, int NV>
__device__ __forceinline__ void block_reduce_all(const T *partials, int stride, int g, T (&tot)[NV]) {
  __shared__ T sh[16 * NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    T a = (T)0;
    for (int i = threadIdx.x; i < g; i += blockDim.x) a += partials[k * stride + i];
    tot[k] = block_sum(a, sh + 16 * k);
  }
}

// Publishes NV partial sums of this workgroup (partials[k*stride + blockIdx])
// and returns true in the workgroup that arrived last; that workgroup then
// finds the totals in tot[0..NV).  `ticket` must start at 0; it is reset by
// the last workgroup, so the next launch can reuse it.
template <typename T, int NV>
__device__ __forceinline__ bool last_block_reduce(const T (&v)[NV], T *partials, int stride,
                                                  unsigned int *ticket, T (&tot)[NV]) {
  __shared__ T sh[16 * NV];
  __shared__ int amlast;
  T bs[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) bs[k] = block_sum(v[k], sh + 16 * k);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      __hip_atomic_store(&partials[k * stride + blockIdx.x], bs[k], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned int t = atomicAdd(ticket, 1u);
    amlast = (t == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!amlast) return false;
  __threadfence();
  const int g = gridDim.x;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    T a = (T)0;
    for (int i = threadIdx.x; i < g; i += blockDim.x) a += load_coherent(&partials[k * stride + i]);
    tot[k] = block_sum(a, sh + 16 * k);
  }
  if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// XCD-aware chunk walk: gridDim.x is a multiple of 8 and the hardware
// dispatches workgroup b to XCD b % 8, so XCD x owns the contiguous chunk
// range [x*C/8, (x+1)*C/8) and its L2 keeps that range's x-vector window.
struct XcdWalk {
  int first, last, step;
  __device__ __forceinline__ XcdWalk(int nchunks) {
    const int xcd = blockIdx.x & 7;
    const int lb = blockIdx.x >> 3;
    const int per = gridDim.x >> 3;
    const long long c0 = (long long)nchunks * xcd / 8;
    const long long c1 = (long long)nchunks * (xcd + 1) / 8;
    first = (int)c0 + lb;
    last = (int)c1;
    step = per;
  }
};

}  // namespace ehip
