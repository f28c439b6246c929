// The no-allocation form of ehip_solve_csr: every device buffer of the
// solve lives in `scratch`, one static device array per GPU, so the call
// makes no hipMalloc/hipFree and creates no stream or event.  The solve is
// the library's FP64 path (solve_impl) on a system held on the host stack
// whose device pointers are carved from that array; the SpMV is the CSR
// stream schedule (its chunk plan is cut on the host in stack-sized pieces).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>

#include "ehip_internal.h"

namespace ehip {

constexpr size_t kScratchBytes = (size_t)EHIP_NOALLOC_SCRATCH_BYTES;
constexpr int kNoallocWork = 11;  // Krylov vectors (BiCGStab's 10) + the unscaled b
constexpr int kNoallocNb = 2048;  // CSR stream chunk (compiled schedule 256 x 2048)

// The library's scratch array on each GPU.  Host code never uses its name
// as an address: it takes the per-device pointer from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch[kMaxDevices];
// host mirrors of the Krylov scalars (read_state copies into them)
static KState g_hst[kMaxDevices];
static KState32 g_hst32[kMaxDevices];

// Device address of the scratch array on the current device, looked up once
// per device with that device current.
static int scratch_ptr(int dev, unsigned char **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("ehip_solve_csr_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      return EHIP_EHIP;
    }
    g_scratch[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch[dev];
  return EHIP_OK;
}

// Consecutive 256-byte aligned pieces of a buffer (base may be null to
// measure a layout).
struct Carve {
  unsigned char *base;
  size_t off = 0;
  template <typename T>
  T *take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T *p = reinterpret_cast<T *>(base + off);
    off += sizeof(T) * count;
    return p;
  }
};

struct NoallocLayout {
  int *rows, *cols, *diag, *chunk;
  double *vals, *rhs, *x, *dscale, *work, *partials;
  float *partials32;
  KState *st;
  KState32 *st32;
  size_t bytes;
};

static NoallocLayout layout(unsigned char *base, size_t n, size_t nnz) {
  Carve c{base};
  NoallocLayout L{};
  L.rows = c.take<int>(n + 1);
  L.cols = c.take<int>(nnz);
  L.diag = c.take<int>(n);
  L.chunk = c.take<int>(n + 1);  // at most one chunk per row
  L.vals = c.take<double>(nnz);
  L.rhs = c.take<double>(n);
  L.x = c.take<double>(n);
  L.dscale = c.take<double>(n);
  L.work = c.take<double>(n * kNoallocWork);
  L.partials = c.take<double>(kPartials);
  L.partials32 = c.take<float>(kPartials);
  L.st = c.take<KState>(1);
  L.st32 = c.take<KState32>(1);
  L.bytes = c.off;
  return L;
}

__global__ void k_to_zero_based(int *__restrict__ a, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    a[i] -= 1;
}

}  // namespace ehip

using namespace ehip;

extern "C" int ehip_solve_csr_noalloc(int n, const int *rows, const int *cols, const int *diag,
                                      const double *values, const double *b, double *x, const ehip_params *p,
                                      ehip_report *r) {
  if (!rows || !cols || !diag || !values || !b || !p || n <= 0) {
    set_error("ehip_solve_csr_noalloc: null argument or n <= 0");
    return EHIP_EARG;
  }
  if (rows[0] != 1 || rows[n] < 1) {
    set_error("ehip_solve_csr_noalloc: rows must be 1-based");
    return EHIP_EARG;
  }
  const int64_t nnz = (int64_t)rows[n] - 1;
  if (n > EHIP_NOALLOC_MAX_N || nnz > EHIP_NOALLOC_MAX_NNZ) {
    set_error("ehip_solve_csr_noalloc: n=%d nnz=%lld exceed the limits n <= %d, nnz <= %d", n, (long long)nnz,
              EHIP_NOALLOC_MAX_N, EHIP_NOALLOC_MAX_NNZ);
    return EHIP_ECAP;
  }
  for (int i = 0; i < n; ++i)
    if (rows[i + 1] < rows[i] || diag[i] < rows[i] || diag[i] >= rows[i + 1] || cols[diag[i] - 1] != i + 1) {
      set_error("ehip_solve_csr_noalloc: row %d is not a sorted CRS row with its diagonal", i + 1);
      return EHIP_EARG;
    }
  for (int64_t j = 0; j < nnz; ++j)
    if (cols[j] < 1 || cols[j] > n) {
      set_error("ehip_solve_csr_noalloc: column index %d out of range at %lld", cols[j], (long long)j + 1);
      return EHIP_EARG;
    }
  if (p->precision != EHIP_FP64 || p->record_history || p->profile_every > 0) {
    set_error("ehip_solve_csr_noalloc: only FP64 solves without history or profiling");
    return EHIP_EUNSUP;
  }
  if (p->precond != EHIP_PRECOND_NONE && p->precond != EHIP_PRECOND_DIAGONAL) {
    // ILU(k) factors and level lists are allocated per system
    set_error("ehip_solve_csr_noalloc: only no preconditioner or the diagonal (ILU allocates its factors)");
    return EHIP_EUNSUP;
  }
  int dev = 0;
  EHIP_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("ehip_solve_csr_noalloc: device %d beyond %d", dev, kMaxDevices);
    return EHIP_EUNSUP;
  }
  unsigned char *base = nullptr;
  int rc = scratch_ptr(dev, &base);
  if (rc) return rc;
  const NoallocLayout L = layout(base, (size_t)n, (size_t)nnz);
  if (L.bytes > kScratchBytes) {  // the limits above keep this from happening
    set_error("ehip_solve_csr_noalloc: %zu bytes of scratch needed, %zu held", L.bytes, kScratchBytes);
    return EHIP_ECAP;
  }
  const hipStream_t st = nullptr;
  // stage the CRS (1-based in, 0-based on the device) and the right-hand side
  EHIP_HIP_TRY(hipMemcpy(L.rows, rows, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice));
  EHIP_HIP_TRY(hipMemcpy(L.cols, cols, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice));
  EHIP_HIP_TRY(hipMemcpy(L.diag, diag, sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
  EHIP_HIP_TRY(hipMemcpy(L.vals, values, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice));
  EHIP_HIP_TRY(hipMemcpy(L.rhs, b, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_to_zero_based, dim3(1024), dim3(256), 0, st, L.rows, (int64_t)n + 1);
  hipLaunchKernelGGL(k_to_zero_based, dim3(1024), dim3(256), 0, st, L.cols, nnz);
  hipLaunchKernelGGL(k_to_zero_based, dim3(1024), dim3(256), 0, st, L.diag, (int64_t)n);
  // the CSR stream chunk plan, cut on the host in pieces of a stack buffer
  const int nplan = chunk_rows(rows, n, kNoallocNb, nullptr, 0, 0);
  {
    int piece[4096];
    for (int first = 0; first < nplan; first += 4096) {
      const int cnt = nplan - first < 4096 ? nplan - first : 4096;
      chunk_rows(rows, n, kNoallocNb, piece, first, cnt);
      EHIP_HIP_TRY(hipMemcpy(L.chunk + first, piece, sizeof(int) * cnt, hipMemcpyHostToDevice));
    }
  }
  // the scalars and their tickets start from zero; everything else the
  // solve writes before it reads
  EHIP_HIP_TRY(hipMemsetAsync(L.st, 0, sizeof(KState), st));
  EHIP_HIP_TRY(hipMemsetAsync(L.st32, 0, sizeof(KState32), st));
  EHIP_HIP_TRY(hipGetLastError());

  ehip_system s;  // host stack; its device buffers are the scratch pieces above
  s.dev = dev;
  s.stream = st;
  s.n = n;
  s.nnz = nnz;
  s.nnodes = n;
  s.rows = L.rows;
  s.cols = L.cols;
  s.diag = L.diag;
  s.vals = L.vals;
  s.rhs = L.rhs;
  s.x = L.x;
  s.dscale = L.dscale;
  s.work = L.work;
  s.nwork = kNoallocWork;
  s.partials = L.partials;
  s.partials32 = L.partials32;
  s.st = L.st;
  s.st32 = L.st32;
  s.h_st = &g_hst[dev];
  s.h_st32 = &g_hst32[dev];
  plan_shape(s.plan64, 256, kNoallocNb, nplan - 1);
  s.plan64.chunk_row = L.chunk;
  s.plan32 = s.plan64;
  s.fmt64 = s.fmt32 = 0;
  s.has_values = true;
  rc = solve_impl(&s, p, x, r, nullptr, false);
  const hipError_t e = hipStreamSynchronize(st);
  if (!rc && e != hipSuccess) {
    set_error("ehip_solve_csr_noalloc: %s", hipGetErrorString(e));
    rc = EHIP_EHIP;
  }
  return rc;
}
