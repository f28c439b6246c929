// Deterministic wave64 / workgroup reductions for the fused Krylov kernels.
// No floating-point atomics and no in-kernel fences: each workgroup publishes
// its partial sums, and a later one-workgroup launch adds them in index
// order, so every result is run-to-run identical.
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
// thread receives the totals.  The partials were written by workgroups on
// every XCD, so each load misses this XCD's L2: a thread issues all of its
// loads (up to kFinPer per value) before adding any, so the reduction costs
// one memory latency instead of one per strided step.
constexpr int kFinPer = 16;
template <typename T, int NV>
__device__ __forceinline__ void block_reduce_all(const T *partials, int stride, int g, T (&tot)[NV]) {
  __shared__ T sh[16 * NV];
  const int nt = blockDim.x;
  if (g <= kFinPer * nt) {
    T v[NV][kFinPer];
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int u = 0; u < kFinPer; ++u) {
        const int i = threadIdx.x + u * nt;  // load clamped (g >= 1), keep only i < g
        const T x = partials[k * stride + min(i, g - 1)];
        v[k][u] = i < g ? x : (T)0;
      }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      T a = (T)0;
#pragma unroll
      for (int u = 0; u < kFinPer; ++u) a += v[k][u];
      tot[k] = block_sum(a, sh + 16 * k);
    }
  } else {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      T a = (T)0;
      for (int i = threadIdx.x; i < g; i += nt) a += partials[k * stride + i];
      tot[k] = block_sum(a, sh + 16 * k);
    }
  }
}

// The same sum as block_reduce_all for g <= PER * blockDim.x, with PER
// loads per thread instead of kFinPer (the dropped terms are zeros, so the
// totals are bit-identical): for kernels whose every workgroup sums a short
// list of partials before its own work.
template <typename T, int NV, int PER>
__device__ __forceinline__ void block_reduce_few(const T *partials, int stride, int g, T (&tot)[NV]) {
  __shared__ T sh[16 * NV];
  const int nt = blockDim.x;
  T v[NV][PER];
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = threadIdx.x + u * nt;  // load clamped (g >= 1), keep only i < g
      const T x = partials[k * stride + min(i, g - 1)];
      v[k][u] = i < g ? x : (T)0;
    }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    T a = (T)0;
#pragma unroll
    for (int u = 0; u < PER; ++u) a += v[k][u];
    tot[k] = block_sum(a, sh + 16 * k);
  }
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
