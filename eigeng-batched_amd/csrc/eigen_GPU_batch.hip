// eigen_GPU_batch.hip - host dispatch and C ABI of the gfx950 batched
// symmetric eigensolver (libeigenGbatch).
//
// Replaces the reference's host layer: eigen_GPU_batch_RUN
// (eigen_GPU_batch.cu:248-325), the C symbols eigen_GPU_batch_DP/_FP
// (eigen_GPU_batch.cu:327-359) and the launch heuristics of
// eigen_GPU_batch_get_Launch_params (eigen_GPU_batch.hpp:23-102).
// No device query is made per call: the launch shape depends only on n and L.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "../../include/eigen_GPU_batch.h"
#include "eig_wave.hpp"
#include "eig_global.hpp"

using namespace eigmi;

namespace {

bool bad_args(int L, int nm, int n, int m) { return L < 1 || nm < n || n < 1 || m < 1 || m > n; }

constexpr int WS_HEAD = 256;      // bytes; keeps the workspace non-empty and aligned

// Which kernel serves (n, element size): register-resident groups while a row
// of the matrix fits in the register file, the workgroup/global-memory kernel
// beyond.
bool use_wave_kernel(int n, int tbytes) { return tbytes == 4 ? n <= 128 : n <= 64; }

int global_grid(int L) { return L < 1024 ? L : 1024; }

// Workspace: the register kernels keep everything in registers and LDS (a
// small fixed block keeps the reference's "wk == NULL means query" protocol
// yielding a non-NULL allocation); the global kernel needs one n x n
// accumulator per resident workgroup.
size_t workspace_bytes(int L, int nm, int n, int m, int tbytes, int solver)
{
  (void)nm; (void)m; (void)solver;
  if (use_wave_kernel(n, tbytes)) return WS_HEAD;
  if (n > GL_NMAX) return 0;
  return WS_HEAD + (size_t)global_grid(L) * (size_t)n * (size_t)n * (size_t)tbytes;
}

template <typename T, int NB>
hipError_t launch_wave(int L, int nm, int n, T *a, T *w, int *info, int solver, hipStream_t s)
{
  constexpr int G = GroupShape<NB>::G, NW = GroupShape<NB>::NW;
  const long long blocks = ((long long)L + G - 1) / G;
  if (blocks > 0x7fffffffll) return hipErrorInvalidValue;
  hipLaunchKernelGGL((eig_wave_kernel<T, NB>), dim3((unsigned)blocks), dim3(64 * NW), 0, s,
                     L, nm, n, a, w, info, solver);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_global(int L, int nm, int n, T *a, T *w, int *info, int solver, void *wk, hipStream_t s)
{
  T *zws = (T *)((char *)wk + WS_HEAD);
  hipLaunchKernelGGL((eig_global_kernel<T>), dim3(global_grid(L)), dim3(GL_THREADS), 0, s,
                     L, nm, n, a, w, info, solver, zws);
  return hipGetLastError();
}

template <typename T>
hipError_t run(int L, int nm, int n, int m, T *a, T *w, void *wk, hipStream_t s, int solver, int *info)
{
  if (bad_args(L, nm, n, m)) return hipErrorInvalidValue;
  if (solver != SOLVER_IMTQL2 && solver != SOLVER_TQL2) return hipErrorNotSupported;
  if (wk == nullptr) return hipErrorInvalidValue;
  if (n <= 8) return launch_wave<T, 8>(L, nm, n, a, w, info, solver, s);
  if (n <= 16) return launch_wave<T, 16>(L, nm, n, a, w, info, solver, s);
  if (n <= 32) return launch_wave<T, 32>(L, nm, n, a, w, info, solver, s);
  if (n <= 64) return launch_wave<T, 64>(L, nm, n, a, w, info, solver, s);
  if (use_wave_kernel(n, (int)sizeof(T))) return launch_wave<T, 128>(L, nm, n, a, w, info, solver, s);
  if (n <= GL_NMAX) return launch_global<T>(L, nm, n, a, w, info, solver, wk, s);
  return hipErrorNotSupported;
}

template <typename T>
void abi_entry(int L, int nm, int n, int m, T *a, T *w, T *wk, hipStream_t s)
{
  if (wk == nullptr) {
    // size query (eigen_GPU_batch.cu:331-339): always a full size_t
    if (w != nullptr) {
      const size_t lw = bad_args(L, nm, n, m) ? 0 : workspace_bytes(L, nm, n, m, (int)sizeof(T), 0);
      memcpy((void *)w, &lw, sizeof(size_t));
    }
    return;
  }
  (void)run<T>(L, nm, n, m, a, w, wk, s, SOLVER_IMTQL2, nullptr);
}

} // namespace

extern "C" {

void eigen_GPU_batch_DP(const int L, const int nm, const int n, const int m, double *a, double *w,
                        double *wk, const gpuStream_t stream)
{
  abi_entry<double>(L, nm, n, m, a, w, wk, stream);
}

void eigen_GPU_batch_FP(const int L, const int nm, const int n, const int m, float *a, float *w,
                        float *wk, const gpuStream_t stream)
{
  abi_entry<float>(L, nm, n, m, a, w, wk, stream);
}

size_t eigen_GPU_batch_bufsize(int L, int nm, int n, int m, int dtype_bytes, int solver)
{
  if (bad_args(L, nm, n, m) || (dtype_bytes != 4 && dtype_bytes != 8)) return 0;
  return workspace_bytes(L, nm, n, m, dtype_bytes, solver);
}

int eigen_GPU_batch_FP_x(int L, int nm, int n, int m, float *a, float *w, void *wk, hipStream_t stream,
                         int solver, int *info)
{
  return (int)run<float>(L, nm, n, m, a, w, wk, stream, solver, info);
}

int eigen_GPU_batch_DP_x(int L, int nm, int n, int m, double *a, double *w, void *wk, hipStream_t stream,
                         int solver, int *info)
{
  return (int)run<double>(L, nm, n, m, a, w, wk, stream, solver, info);
}

const char *eigen_GPU_batch_version(void) { return "eigen-mi355x 0.1"; }

} // extern "C"
