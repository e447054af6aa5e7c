// eig_symdmt.hpp - the dense batched products on either side of the solver, on
// MFMA (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64, FP32 / FP64 accumulate):
//
//   eig_symdmt_kernel  C_k = Z_k diag(d_k) Z_k^T, with d_k given or
//                      d_k = scale * w_k^p (the LETKF consumer step: the analysis
//                      covariance V L^-1 V^T, the transform sqrt(k-1) V L^-1/2 V^T).
//                      Reference: the batched symdmt_batch experiment (symdmt.cu:5-87)
//                      and "any power of the input matrix" (README:171).
//   eig_gram_kernel    A_k = alpha I + Y_k^T diag(r_k) Y_k, Y_k of p x n (the
//                      LETKF ingest that forms the matrices the solver takes;
//                      the reference forms its inputs on the host, main.cpp:63-195).
//
// Both outputs are symmetric: a workgroup owns one matrix and its waves take the
// 32 x 32 output blocks (I, J), I <= J, in turn; a block is 2 x 2 MFMA tiles of
// 16 x 16 accumulated over the contraction in steps of 4.  Every block value is
// written to both (i, j) and (j, i), so the result is exactly symmetric.
//
// MFMA operand layout (16x16x4, wave64): lane l = m + 16 g holds A(m, g) and
// B(g, m); accumulator slot r of lane l holds C(row(g, r), m) with
// row = 4 g + r for f32 and g + 4 r for f64.
#pragma once
#include "common.hpp"
#include "eig_global.hpp"

namespace eigmi {

template <typename T> struct Mfma16;
template <> struct Mfma16<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};
template <> struct Mfma16<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};

constexpr int SD_MAX_THREADS = 256;

// block index -> (I, J), I <= J, blocks ordered by J then I
__device__ __forceinline__ void sym_block(int blk, int &I, int &J)
{
  J = 0;
  while ((J + 1) * (J + 2) / 2 <= blk) ++J;
  I = blk - J * (J + 1) / 2;
}

// Write the 32 x 32 block (I, J) held in acc to (i, j) and (j, i) of the
// column-major O (leading dimension nm); diag is added on i == j.
template <typename T>
__device__ __forceinline__ void store_sym_block(T *__restrict__ O, int nm, int n, int I, int J,
                                                const typename Mfma16<T>::v4 (&acc)[2][2], T diag)
{
  const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 32 * I + 16 * a + Mfma16<T>::row(g, r);
        const int j = 32 * J + 16 * b + m;
        if (i < n && j < n && i <= j) {
          const T v = acc[a][b][r] + ((i == j) ? diag : T(0));
          O[j + (size_t)i * nm] = v;      // lanes m: consecutive rows of column i
          O[i + (size_t)j * nm] = v;
        }
      }
}

// d_k from the source row: fn 0: as given; fn 1: scale * w^p
template <typename T>
__device__ __forceinline__ T dfun(T w, int fn, T p, T scale)
{
  if (fn == 0) return w;
  T v;
  if (p == T(1)) v = w;
  else if (p == T(-1)) v = T(1) / w;
  else if (p == T(0.5)) v = sqrt(w);
  else if (p == T(-0.5)) v = T(1) / sqrt(w);
  else if (p == T(2)) v = w * w;
  else if (p == T(0)) v = T(1);
  else v = pow(w, p);
  return scale * v;
}

template <typename T>
__global__ __launch_bounds__(SD_MAX_THREADS) void eig_symdmt_kernel(int L, int nm, int n, const T *__restrict__ z,
                                                                    const T *__restrict__ dsrc, int fn, T p,
                                                                    T scale, T *__restrict__ out)
{
  typedef typename Mfma16<T>::v4 v4;
  __shared__ T ds[GL_NMAX];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, nw = blockDim.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int nb = (n + 31) >> 5, nblk = nb * (nb + 1) / 2;
  for (long long mat = blockIdx.x; mat < L; mat += gridDim.x) {
    const T *__restrict__ Z = z + (size_t)mat * nm * n;
    const T *__restrict__ D = dsrc + (size_t)mat * n;
    T *__restrict__ O = out + (size_t)mat * nm * n;
    for (int i = t; i < n; i += blockDim.x) ds[i] = dfun(D[i], fn, p, scale);
    __syncthreads();
    for (int blk = wv; blk < nblk; blk += nw) {
      int I, J;
      sym_block(blk, I, J);
      // operand rows (clamped into the matrix; rows >= n only feed outputs
      // that are not stored)
      const int ra0 = min(32 * I + m, n - 1), ra1 = min(32 * I + 16 + m, n - 1);
      const int rb0 = min(32 * J + m, n - 1), rb1 = min(32 * J + 16 + m, n - 1);
      v4 acc[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4{0, 0, 0, 0};
#pragma unroll 4
      for (int k0 = 0; k0 < n; k0 += 4) {
        const int k = k0 + g;
        const bool kv = k < n;
        const T *__restrict__ col = Z + (size_t)(kv ? k : n - 1) * nm;
        const T dk = kv ? ds[k] : T(0);
        const T a0 = kv ? col[ra0] * dk : T(0), a1 = kv ? col[ra1] * dk : T(0);
        const T b0 = col[rb0], b1 = col[rb1];
        acc[0][0] = Mfma16<T>::op(a0, b0, acc[0][0]);
        acc[0][1] = Mfma16<T>::op(a0, b1, acc[0][1]);
        acc[1][0] = Mfma16<T>::op(a1, b0, acc[1][0]);
        acc[1][1] = Mfma16<T>::op(a1, b1, acc[1][1]);
      }
      store_sym_block<T>(O, nm, n, I, J, acc, T(0));
    }
    __syncthreads();
  }
}

// Four consecutive elements p[0..3] (those at index >= cnt read as 0).  VEC:
// one 16-byte (FP32) or two 16-byte (FP64) loads; the caller guarantees
// alignment and that cnt is a multiple of 4 (so cnt > 0 means cnt >= 4).
template <typename T, bool VEC>
__device__ __forceinline__ void load4(const T *__restrict__ p, int cnt, T (&v)[4])
{
  if constexpr (VEC) {
    if constexpr (sizeof(T) == 4) {
      const float4 x = *reinterpret_cast<const float4 *>(p);
      v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
    } else {
      const double2 x = reinterpret_cast<const double2 *>(p)[0], y = reinterpret_cast<const double2 *>(p)[1];
      v[0] = x.x; v[1] = x.y; v[2] = y.x; v[3] = y.y;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = (cnt > 0) ? v[q] : T(0);   // cnt is 0 mod 4 here
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = (q < cnt) ? p[q < cnt ? q : 0] : T(0);
  }
}

// A_k = alpha I + Y_k^T diag(r_k) Y_k.  Y_k: p x n, column-major, leading
// dimension ldy, batch stride ldy * n; r_k = rinv + k * rstride (rinv null:
// r = 1; rstride 0: one vector for the batch).  The contraction index o runs
// over 16 per step: lane (m, g) loads o = o0 + 4 g + q, q = 0..3, of its
// columns (16 contiguous bytes per lane and column), and MFMA q of the step
// takes element q, so A- and B-operands agree on the order of o.
// VEC requires ldy % 4 == 0, p % 4 == 0 and a 16-byte aligned y (and rinv).
template <typename T, bool VEC>
__global__ __launch_bounds__(SD_MAX_THREADS) void eig_gram_kernel(int L, int nm, int n, int np,
                                                                  const T *__restrict__ y, int ldy,
                                                                  const T *__restrict__ rinv, int rstride, T alpha,
                                                                  T *__restrict__ out)
{
  typedef typename Mfma16<T>::v4 v4;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, nw = blockDim.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int nb = (n + 31) >> 5, nblk = nb * (nb + 1) / 2;
  for (long long mat = blockIdx.x; mat < L; mat += gridDim.x) {
    const T *__restrict__ Y = y + (size_t)mat * ldy * n;
    const T *__restrict__ R = rinv ? rinv + (size_t)mat * rstride : nullptr;
    T *__restrict__ O = out + (size_t)mat * nm * n;
    for (int blk = wv; blk < nblk; blk += nw) {
      int I, J;
      sym_block(blk, I, J);
      const T *__restrict__ ya0 = Y + (size_t)min(32 * I + m, n - 1) * ldy;
      const T *__restrict__ ya1 = Y + (size_t)min(32 * I + 16 + m, n - 1) * ldy;
      const T *__restrict__ yb0 = Y + (size_t)min(32 * J + m, n - 1) * ldy;
      const T *__restrict__ yb1 = Y + (size_t)min(32 * J + 16 + m, n - 1) * ldy;
      v4 acc[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4{0, 0, 0, 0};
#pragma unroll 2
      for (int o0 = 0; o0 < np; o0 += 16) {
        const int o = o0 + 4 * g;
        const int cnt = np - o;                 // valid elements from o on (may be <= 0)
        const int oc = cnt > 0 ? o : 0;         // clamped start when none is valid
        T xa0[4], xa1[4], xb0[4], xb1[4], rv[4];
        load4<T, VEC>(ya0 + oc, cnt, xa0);
        load4<T, VEC>(ya1 + oc, cnt, xa1);
        load4<T, VEC>(yb0 + oc, cnt, xb0);
        load4<T, VEC>(yb1 + oc, cnt, xb1);
        if (R) {
          load4<T, VEC>(R + oc, cnt, rv);
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) rv[q] = (q < cnt) ? T(1) : T(0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const T a0 = xa0[q] * rv[q], a1 = xa1[q] * rv[q];
          acc[0][0] = Mfma16<T>::op(a0, xb0[q], acc[0][0]);
          acc[0][1] = Mfma16<T>::op(a0, xb1[q], acc[0][1]);
          acc[1][0] = Mfma16<T>::op(a1, xb0[q], acc[1][0]);
          acc[1][1] = Mfma16<T>::op(a1, xb1[q], acc[1][1]);
        }
      }
      store_sym_block<T>(O, nm, n, I, J, acc, alpha);
    }
  }
}

// ---- n <= 64: one wave per matrix, operands held in registers ----------------
// The 16-row tiles t = 0..NT-1 of the matrix; the wave forms the tile pairs
// (ti, tj), ti <= tj, so each operand element is loaded once per matrix.
template <typename T, int NT>
__device__ __forceinline__ void store_sym_tiles(T *__restrict__ O, int nm, int n, int m, int g,
                                                const typename Mfma16<T>::v4 (&acc)[NT * (NT + 1) / 2], T diag)
{
  int idx = 0;
#pragma unroll
  for (int ti = 0; ti < NT; ++ti)
#pragma unroll
    for (int tj = ti; tj < NT; ++tj, ++idx)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * ti + Mfma16<T>::row(g, r);
        const int j = 16 * tj + m;
        if (i < n && j < n && i <= j) {
          const T v = acc[idx][r] + ((i == j) ? diag : T(0));
          O[j + (size_t)i * nm] = v;
          O[i + (size_t)j * nm] = v;
        }
      }
}

// occupancy floor for the register-resident kernels (waves per SIMD)
template <typename T> struct SdWaves { static constexpr int value = sizeof(T) == 8 ? 2 : 3; };

template <typename T, int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SdWaves<T>::value))) void eig_symdmt_wave_kernel(int L, int nm, int n, const T *__restrict__ z,
                                                              const T *__restrict__ dsrc, int fn, T p, T scale,
                                                              T *__restrict__ out)
{
  typedef typename Mfma16<T>::v4 v4;
  constexpr int KS = 4 * NT, NP = NT * (NT + 1) / 2;     // k-steps, tile pairs
  constexpr int KC = (sizeof(T) == 8 && NT >= 3) ? 8 : KS;
  const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long mat = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); mat < L; mat += nw) {
    // lane indices made opaque per matrix: otherwise every lane-constant
    // address and mask is hoisted out of this loop into its own register
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    const int m = lane & 15, g = lane >> 4;
    const T *__restrict__ Z = z + (size_t)mat * nm * n;
    const T *__restrict__ D = dsrc + (size_t)mat * n;
    int rows[NT];                                           // operand rows, clamped into the matrix
#pragma unroll
    for (int t = 0; t < NT; ++t) rows[t] = min(16 * t + m, n - 1);
    v4 acc[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[q] = v4{0, 0, 0, 0};
    // KC k-steps at a time (all of them for FP32; FP64 keeps half the operands)
#pragma unroll 1
    for (int k0 = 0; k0 < KS; k0 += KC) {
      T R[NT][KC], dk[KC];
#pragma unroll
      for (int kk = 0; kk < KC; ++kk) {
        const int kc = min(4 * (k0 + kk) + g, n - 1);         // clamped column of the step
        const T *__restrict__ Zk = Z + (size_t)kc * nm;
#pragma unroll
        for (int t = 0; t < NT; ++t) R[t][kk] = Zk[rows[t]];
        dk[kk] = D[kc];
      }
#pragma unroll
      for (int kk = 0; kk < KC; ++kk)
        dk[kk] = (4 * (k0 + kk) + g < n) ? dfun(dk[kk], fn, p, scale) : T(0);
#pragma unroll
      for (int kk = 0; kk < KC; ++kk) {
        // (steps beyond n carry dk = 0 and add nothing)
#pragma unroll
        for (int ti = 0, idx = 0; ti < NT; ++ti) {
          const T a = R[ti][kk] * dk[kk];
#pragma unroll
          for (int tj = ti; tj < NT; ++tj, ++idx) acc[idx] = Mfma16<T>::op(a, R[tj][kk], acc[idx]);
        }
      }
    }
    store_sym_tiles<T, NT>(out + (size_t)mat * nm * n, nm, n, m, g, acc, T(0));
  }
}

template <typename T, int NT, bool VEC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SdWaves<T>::value))) void eig_gram_wave_kernel(int L, int nm, int n, int np, const T *__restrict__ y,
                                                            int ldy, const T *__restrict__ rinv, int rstride,
                                                            T alpha, T *__restrict__ out)
{
  typedef typename Mfma16<T>::v4 v4;
  constexpr int NP = NT * (NT + 1) / 2;
  const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long mat = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); mat < L; mat += nw) {
    int lane = threadIdx.x & 63;                           // opaque per matrix (see above)
    asm volatile("" : "+v"(lane));
    const int m = lane & 15, g = lane >> 4;
    const T *__restrict__ Y = y + (size_t)mat * ldy * n;
    const T *__restrict__ R = rinv ? rinv + (size_t)mat * rstride : nullptr;
    const T *yc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) yc[t] = Y + (size_t)min(16 * t + m, n - 1) * ldy;
    v4 acc[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[q] = v4{0, 0, 0, 0};
    // software pipeline: the next step's loads are issued before this step's MFMAs
    T x[NT][4], rv[4], xn[NT][4] = {}, rn[4] = {};
    auto fetch = [&](int o0, T (&xo)[NT][4], T (&ro)[4]) {
      const int o = o0 + 4 * g, cnt = np - o, oc = cnt > 0 ? o : 0;
#pragma unroll
      for (int t = 0; t < NT; ++t) load4<T, VEC>(yc[t] + oc, cnt, xo[t]);
      if (R) {
        load4<T, VEC>(R + oc, cnt, ro);
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) ro[q] = (q < cnt) ? T(1) : T(0);
      }
    };
    fetch(0, x, rv);
    for (int o0 = 0; o0 < np; o0 += 16) {
      if (o0 + 16 < np) fetch(o0 + 16, xn, rn);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int idx = 0;
#pragma unroll
        for (int ti = 0; ti < NT; ++ti) {
          const T a = x[ti][q] * rv[q];
#pragma unroll
          for (int tj = ti; tj < NT; ++tj, ++idx) acc[idx] = Mfma16<T>::op(a, x[tj][q], acc[idx]);
        }
      }
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) x[t][q] = xn[t][q];
#pragma unroll
      for (int q = 0; q < 4; ++q) rv[q] = rn[q];
    }
    store_sym_tiles<T, NT>(out + (size_t)mat * nm * n, nm, n, m, g, acc, alpha);
  }
}

} // namespace eigmi
