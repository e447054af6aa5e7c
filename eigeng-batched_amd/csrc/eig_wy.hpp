// eig_wy.hpp - blocked (compact WY) back-transformation on the matrix cores:
// V = H_{n-1} ... H_1 Z for the reflectors H_c = I + beta_c u_c u_c^T that the
// large-order tridiagonalisation leaves in its work array (u_c in rows 0..c-1
// of column c), applied to the tridiagonal's eigenvectors Z.  The reference
// aggregates its reflectors in blocks of 3-5 and reduces the products with
// shuffles (hhtr2sy.hpp:69-203, hhtr2sy_tiled.hpp:82-211); here the blocks are
// 64 reflectors wide and every product is a batched matrix-core GEMM:
//
//   block k = reflectors c0 .. c0+nb-1 (applied in increasing c):
//   H_{c0+nb-1} ... H_{c0} = I + U T U^T,  T lower triangular,
//   T_jj = beta_j,  T_j,0:j = beta_j (u_j^T U_0:j) T_0:j,0:j   (adding H_j on the left)
//   V <- V + U (T (U^T V))
//
// The u_c are first scaled to unit length (beta_c ||u_c||^2 replaces beta_c,
// which is -2 for an orthogonal reflector), so every operand is bounded and the
// FP16-split products (GK_F16X3) stay in range whatever the matrix's scale.
#pragma once
#include "common.hpp"
#include "eig_wave.hpp"

namespace eigmi {

constexpr int WY_NB = 64;

// One wave per (matrix, column c): u_c <- u_c / ||u_c|| over rows 0..c-1, rows
// c..n-1 of column c <- 0, bwy_c = beta_c ||u_c||^2 (0 for c = 0, a zero u_c,
// or a matrix whose status is set: those keep V as it is).
template <typename T>
__global__ __launch_bounds__(256) void wy_prepare_kernel(int L, int n, T *__restrict__ W, const T *__restrict__ bw,
                                                         T *__restrict__ bwy, const int *__restrict__ status)
{
  const int lane = threadIdx.x & 63;
  const long long cols = (long long)L * n;
  for (long long q = (long long)blockIdx.x * 4 + (threadIdx.x >> 6); q < cols; q += (long long)gridDim.x * 4) {
    const long long b = q / n;
    const int c = (int)(q - b * n);
    T *__restrict__ u = W + (size_t)b * n * n + (size_t)c * n;
    const bool live = status[b] == 0 && c >= 1;
    T ss = T(0);
    for (int r = lane; r < c; r += 64) ss = fma_(u[r], u[r], ss);
    ss = group_sum<64>(ss);
    const bool ok = live && ss > T(0);
    const T inv = ok ? T(1) / sqrt_(ss) : T(0);
    for (int r = lane; r < n; r += 64) u[r] = (ok && r < c) ? u[r] * inv : T(0);
    if (lane == 0) bwy[(size_t)b * n + c] = ok ? bw[(size_t)b * n + c] * ss : T(0);
  }
}

// T of block (c0, nb) from G = U^T U (ld WY_NB) and the scaled betas, one wave
// per matrix, lane i computing column i (T in LDS, row by row).
template <typename T>
__global__ __launch_bounds__(64) void wy_t_kernel(int L, int n, int c0, int nb, const T *__restrict__ Gm,
                                                  const T *__restrict__ bwy, T *__restrict__ Tm)
{
  __shared__ T Gs[WY_NB][WY_NB + 1];
  __shared__ T Ts[WY_NB][WY_NB + 1];
  const int i = threadIdx.x;
  for (long long b = blockIdx.x; b < L; b += gridDim.x) {
    const T *__restrict__ G = Gm + (size_t)b * WY_NB * WY_NB;
    for (int j = 0; j < nb; ++j) Gs[j][i] = (i < nb) ? G[i + (size_t)j * WY_NB] : T(0);   // Gs[col][row]
    wsync();
    for (int j = 0; j < nb; ++j) {
      const T bj = bwy[(size_t)b * n + c0 + j];
      T v = T(0);
      if (i < j) {
        T s = T(0);
        for (int l = i; l < j; ++l) s = fma_(Gs[l][j], Ts[l][i], s);   // G[j][l] = u_j^T u_l
        v = bj * s;
      } else if (i == j) {
        v = bj;
      }
      Ts[j][i] = v;
      wsync();
    }
    T *__restrict__ Tb = Tm + (size_t)b * WY_NB * WY_NB;
    for (int j = 0; j < WY_NB; ++j) Tb[j + (size_t)i * WY_NB] = (i < nb && j < nb) ? Ts[j][i] : T(0);
    wsync();
  }
}

} // namespace eigmi
