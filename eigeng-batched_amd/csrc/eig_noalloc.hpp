// eig_noalloc.hpp - the no-allocation entry points of libeigenGbatch.
//
// Included at the end of eigen_GPU_batch.hip (it uses that translation unit's
// workspace sizing and dispatch).  No counterpart in the reference: for
// callers that cannot allocate device memory.
#pragma once

// A fixed device array serves as the workspace for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Each device that loads the library holds its own copy
// until the process ends.  Host code never passes `scratch` itself to a
// kernel (in host code the name is a host-side stand-in): its device address
// is looked up once per device with hipGetSymbolAddress and kept below.
static constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

hipError_t scratch_ptr(void **out)
{
  static std::mutex mu;
  static std::map<int, void *> addr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(mu);
  auto it = addr.find(dev);
  if (it != addr.end()) { *out = it->second; return hipSuccess; }
  void *ptr = nullptr;
  e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (ptr == nullptr) return hipErrorInvalidSymbol;
  addr[dev] = ptr;
  *out = ptr;
  return hipSuccess;
}

// The _x computation with the workspace in `scratch`.  A batch whose workspace
// exceeds the array is solved in consecutive sub-batches (halving the count
// until it fits).  None of the kernels needs the workspace cleared: every
// region is written before it is read within a call.
template <typename T>
hipError_t run_noalloc(int L, int nm, int n, int m, T *a, T *w, hipStream_t s, int solver, int *info)
{
  if (bad_args(L, nm, n, m) || !a || !w) return hipErrorInvalidValue;
  if (solver & (FLAG_PH_ALL | FLAG_TINYLOG)) return hipErrorInvalidValue;   // phases hand over through wk
  if (n > GL_NMAX) return hipErrorNotSupported;
  const int tb = (int)sizeof(T);
  int lc = L;
  while (workspace_bytes(lc, nm, n, m, tb, solver) > NOALLOC_SCRATCH_BYTES && lc > 1) lc = (lc + 1) / 2;
  if (workspace_bytes(lc, nm, n, m, tb, solver) > NOALLOC_SCRATCH_BYTES) {
    fprintf(stderr, "eigen_GPU_batch_noalloc: one matrix of order %d needs %zu workspace bytes, more than the "
            "%zu of the fixed scratch array\n", n, workspace_bytes(1, nm, n, m, tb, solver),
            NOALLOC_SCRATCH_BYTES);
    return hipErrorInvalidValue;
  }
  void *ws = nullptr;
  hipError_t e = scratch_ptr(&ws);
  if (e != hipSuccess) return e;
  for (long long c0 = 0; c0 < L; c0 += lc) {
    const int cur = (int)((L - c0) < lc ? (L - c0) : lc);
    e = run<T>(cur, nm, n, m, a + (size_t)c0 * nm * n, w + (size_t)c0 * n, ws, s, solver,
               info ? info + c0 : nullptr);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

} // namespace

extern "C" {

size_t eigen_GPU_batch_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

int eigen_GPU_batch_FP_noalloc(int L, int nm, int n, int m, float *a, float *w, hipStream_t stream, int solver,
                               int *info)
{
  return (int)run_noalloc<float>(L, nm, n, m, a, w, stream, solver, info);
}

int eigen_GPU_batch_DP_noalloc(int L, int nm, int n, int m, double *a, double *w, hipStream_t stream, int solver,
                               int *info)
{
  return (int)run_noalloc<double>(L, nm, n, m, a, w, stream, solver, info);
}

} // extern "C"
