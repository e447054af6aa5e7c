// eig_uneven.hpp - batches of matrices of different orders in one call.
//
// The reference solves one order n per call (eigen_GPU_batch.cu:248-325); a
// batch of mixed orders costs one call per order there.  Here the orders stay
// on the device:
//   1. uv_bin_kernel (one workgroup) counts the orders, lays the matrices out
//      in a list sorted by descending order (counting sort, each order's run
//      aligned to the matrices a wave takes at once), records for every
//      size class the list range [begin, end) and a work cursor, and, for the
//      packed layout, forms the element offsets of every matrix and of its
//      eigenvalues by prefix sums;
//   2. one persistent launch per size class (groups of 8/16/32/64/128 lanes or
//      the workgroup kernel) takes work from the class's cursor, largest
//      matrices first, until the class's range is drained.
// Every step is a kernel (no memset, no host synchronisation, no allocation),
// so the whole sequence can be captured in a HIP graph.
#pragma once
#include "common.hpp"
#include "eig_wave.hpp"
#include "eig_global.hpp"
#include "eig_split.hpp"

namespace eigmi {

constexpr int UV_MAXCLS = 8;
constexpr int UV_BAD_ORDER = -3;  // info value of a matrix whose order is out of range
constexpr int UV_THREADS = 1024;

// Size classes: class c serves orders (hi[c-1], hi[c]] (hi[-1] = 0).
// Every order's run in the sorted list starts at a multiple of its class's
// matrices-per-wave g[c], so the g entries a wave takes at once share one
// order (which keeps that order wave-uniform); the gaps hold -1.
struct UvClasses {
  int ncls;
  int hi[UV_MAXCLS];
  int g[UV_MAXCLS];
};

// List length for L matrices: room for the alignment gaps.
__host__ __device__ constexpr long long uv_list_len(int L) { return (long long)L + 8LL * (GL_NMAX + 1); }

// ctl layout per class: [begin, end, cursor, unused]
__global__ __launch_bounds__(UV_THREADS) void uv_bin_kernel(int L, int nmax, const int *__restrict__ nv,
                                                            long long *__restrict__ aoff_ws,
                                                            long long *__restrict__ woff_ws, int packed,
                                                            int *__restrict__ list, int *__restrict__ ctl,
                                                            int *__restrict__ info, UvClasses cls)
{
  __shared__ int hist[GL_NMAX + 2];
  __shared__ int cur[GL_NMAX + 2];
  __shared__ long long psa[UV_THREADS], psw[UV_THREADS];
  const int t = threadIdx.x;
  for (int i = t; i < GL_NMAX + 2; i += UV_THREADS) hist[i] = 0;
  __syncthreads();
  for (int k = t; k < L; k += UV_THREADS) {
    const int nk = nv[k];
    if (nk >= 1 && nk <= nmax) atomicAdd(&hist[nk], 1);
    else if (info) info[k] = UV_BAD_ORDER;
  }
  __syncthreads();
  const int llen = (int)uv_list_len(L);
  for (int i = t; i < llen; i += UV_THREADS) list[i] = -1;
  if (t == 0) {
    // descending counting-sort positions, each run aligned to its class's g
    int pos = 0, c = cls.ncls - 1;
    for (int n = nmax; n >= 1; --n) {
      while (c > 0 && cls.hi[c - 1] >= n) --c;
      const int g = cls.g[c];
      pos = (pos + g - 1) / g * g;
      cur[n] = pos;
      pos += hist[n];
    }
    for (int c = 0; c < cls.ncls; ++c) {
      const int lo = (c == 0) ? 0 : cls.hi[c - 1];
      const int hi = cls.hi[c] < nmax ? cls.hi[c] : nmax;
      int b = 0, e = 0;
      if (hi > lo) { b = cur[hi]; e = cur[lo + 1] + hist[lo + 1]; }
      ctl[4 * c + 0] = b;
      ctl[4 * c + 1] = e;
      ctl[4 * c + 2] = b;
      ctl[4 * c + 3] = 0;
    }
  }
  __syncthreads();
  for (int k = t; k < L; k += UV_THREADS) {
    const int nk = nv[k];
    if (nk >= 1 && nk <= nmax) list[atomicAdd(&cur[nk], 1)] = k;
  }
  if (!packed) return;
  // packed layout: exclusive prefix sums of n^2 and n in matrix order
  const int chunk = (L + UV_THREADS - 1) / UV_THREADS;
  const int k0 = t * chunk, k1 = (k0 + chunk < L) ? k0 + chunk : L;
  long long sa = 0, sw = 0;
  for (int k = k0; k < k1; ++k) {
    const long long nk = nv[k] > 0 ? nv[k] : 0;
    sa += nk * nk;
    sw += nk;
  }
  psa[t] = sa;
  psw[t] = sw;
  __syncthreads();
  for (int off = 1; off < UV_THREADS; off <<= 1) {
    const long long xa = (t >= off) ? psa[t - off] : 0, xw = (t >= off) ? psw[t - off] : 0;
    __syncthreads();
    psa[t] += xa;
    psw[t] += xw;
    __syncthreads();
  }
  sa = psa[t] - sa;
  sw = psw[t] - sw;
  for (int k = k0; k < k1; ++k) {
    aoff_ws[k] = sa;
    woff_ws[k] = sw;
    const long long nk = nv[k] > 0 ? nv[k] : 0;
    sa += nk * nk;
    sw += nk;
  }
}

// Waves per SIMD the one-order kernel reaches for (T, NB); the list kernel is
// held to the same register budget.
template <typename T, int NB> struct UvWaves { static constexpr int value = 1; };
template <> struct UvWaves<float, 8> { static constexpr int value = 6; };
template <> struct UvWaves<float, 16> { static constexpr int value = 4; };
template <> struct UvWaves<float, 32> { static constexpr int value = 3; };
template <> struct UvWaves<float, 64> { static constexpr int value = 3; };
template <> struct UvWaves<double, 8> { static constexpr int value = 4; };
template <> struct UvWaves<double, 16> { static constexpr int value = 2; };

// Register-resident groups of NB lanes; each wave (or two-wave workgroup)
// takes G list entries at a time from the class cursor.
template <typename T, int NB>
__global__ __launch_bounds__(64 * GroupShape<NB>::NW)
__attribute__((amdgpu_waves_per_eu(UvWaves<T, NB>::value))) void eig_wave_list_kernel(
    int *__restrict__ ctl, const int *__restrict__ list, const int *__restrict__ nv,
    const long long *__restrict__ aoff, const long long *__restrict__ woff, T *__restrict__ a,
    T *__restrict__ w, int *__restrict__ info, int solver)
{
  constexpr int NL = GroupShape<NB>::NL, G = GroupShape<NB>::G;
  __shared__ WaveLds<T, NB> smem[G];
  __shared__ int take;
  const int lane = threadIdx.x & 63;
  const int wv = (int)(threadIdx.x >> 6);
  const int g = lane / NL, kl = lane % NL;
  const int k = wv * NL + kl;
  const int end = ctl[1];
  for (;;) {
    if (threadIdx.x == 0) take = atomicAdd(ctl + 2, G);
    __syncthreads();
    const int t0 = __builtin_amdgcn_readfirstlane(take);
    __syncthreads();
    if (t0 >= end) break;
    const int t = t0 + g;
    // the first entry of an aligned run is a matrix, and all G share its order
    const int n = __builtin_amdgcn_readfirstlane(nv[list[t0]]);
    int mat = (t < end) ? list[t] : -1;
    if (G == 1) mat = __builtin_amdgcn_readfirstlane(mat);
    // lane indices made opaque per matrix: keeps the compiler from hoisting
    // values derived from them (identity rows, masks) out of this loop, where
    // they would stay live in registers across the whole solve
    int lane_i = lane, kl_i = kl, k_i = k;
    asm volatile("" : "+v"(lane_i), "+v"(kl_i), "+v"(k_i));
    if (mat >= 0)
      wave_solve<T, NB>(a + aoff[mat], n, n, w + woff[mat], info ? info + mat : nullptr, solver, smem[g],
                        lane_i, wv, kl_i, k_i);
  }
}

// Workgroup kernel for the orders beyond the register kernels; Z is one
// zstride-element accumulator per workgroup.
template <typename T>
__global__ __launch_bounds__(GL_THREADS) void eig_global_list_kernel(
    int *__restrict__ ctl, const int *__restrict__ list, const int *__restrict__ nv,
    const long long *__restrict__ aoff, const long long *__restrict__ woff, T *__restrict__ a,
    T *__restrict__ w, int *__restrict__ info, int solver, T *__restrict__ zws, long long zstride)
{
  __shared__ GlobalLds<T> S;
  __shared__ int take;
  T *__restrict__ Z = zws + (size_t)blockIdx.x * (size_t)zstride;
  const int end = ctl[1];
  for (;;) {
    if (threadIdx.x == 0) take = atomicAdd(ctl + 2, 1);
    __syncthreads();
    const int t0 = __builtin_amdgcn_readfirstlane(take);
    __syncthreads();
    if (t0 >= end) break;
    const int mat = list[t0];
    const int n = __builtin_amdgcn_readfirstlane(nv[mat]);
    global_solve<T, GL_THREADS>(a + aoff[mat], n, n, w + woff[mat], info ? info + mat : nullptr, solver, Z, S);
  }
}

// Split route for a size class: when the class has at least min_count
// entries, move its work cursor to the end, so that the fused list kernel of
// the class finds nothing left to take.
__global__ void uv_claim_kernel(int *__restrict__ ctl, int min_count)
{
  if (threadIdx.x == 0 && blockIdx.x == 0 && ctl[1] - ctl[0] >= min_count) ctl[2] = ctl[1];
}

// Split route, per chunk: the entries whose rotation log overflowed (status
// ST_OVERFLOW) are solved from their untouched input by the workgroup kernel.
template <typename T>
__global__ __launch_bounds__(GL_THREADS) void eig_global_fallback_list_kernel(UvSplitArgs u,
                                                                              const int *__restrict__ status,
                                                                              T *__restrict__ a, T *__restrict__ w,
                                                                              int *__restrict__ info, int solver,
                                                                              T *__restrict__ zws)
{
  __shared__ GlobalLds<T> S;
  T *__restrict__ Z = zws + (size_t)blockIdx.x * u.nmx * u.nmx;
  for (long long idx = blockIdx.x; idx < u.cap; idx += gridDim.x) {
    int mat;
    if (!uv_split_entry(u, idx, mat)) continue;        // the same for the whole workgroup
    if (status[idx] != ST_OVERFLOW) continue;
    const int n = __builtin_amdgcn_readfirstlane(u.nv[mat]);
    global_solve<T, GL_THREADS>(a + u.aoff[mat], n, n, w + u.woff[mat], info ? info + mat : nullptr, solver, Z, S);
  }
}

} // namespace eigmi
