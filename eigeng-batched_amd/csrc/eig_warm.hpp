// eig_warm.hpp - the warm-started solve for n <= 64: an approximate eigenbasis
// Q0 (e.g. predicted by the warm-start network) is made orthonormal by a
// batched Householder QR, the matrix is rotated into it (B = Q^T A Q, matrix
// cores), a few cyclic Jacobi sweeps finish B (eig_jacobi_kernel with a sweep
// cap), and V = Q Z_B (matrix cores).  A matrix whose Jacobi needs more sweeps
// than the cap is solved by the plain solver (tridiagonalisation + QL,
// eig_wave_kernel with `only`) from its untouched input.
//
// The reference's Jacobi (jac_tiled.hpp:3-152) always starts from the identity;
// starting from Q0 is the same iteration on Q0^T A Q0.
#pragma once
#include "common.hpp"
#include "eig_wave.hpp"

namespace eigmi {

// ---- batched Householder QR (Q factor), n <= 64, one matrix per wave ----------
// Lane k owns row k of the matrix in LDS (m[j][k] = element (k, j), rows padded
// to 65 so that a lane walking its own column and the lanes of a row are both
// conflict free).  Column c: alpha = -sign(x_c) ||x||, v = x - alpha e_c,
// tau = 2 / v^T v; the trailing columns take w_j = v^T m[:, j] (lane j) and
// m[:, j] -= tau v w_j (lane per row).  Q = H_0 ... H_{n-2} is accumulated
// backwards from the identity, touching only its trailing block, and its
// column j is negated where R_jj < 0, so R has a non-negative diagonal (the
// factor of a matrix with positive-diagonal R is unique).
template <typename T>
struct QrLds {
  T m[64][65];
  T q[64][65];
  T v[64];
  T w[64];
  T tau[64];
  T sg[64];
};

template <typename T>
__global__ __launch_bounds__(64) void eig_qr_kernel(int L, int nm, int n, const T *__restrict__ a,
                                                    T *__restrict__ qout, long long ldq, long long sq)
{
  __shared__ QrLds<T> S;
  const int k = threadIdx.x;
  for (long long mat = blockIdx.x; mat < L; mat += gridDim.x) {
    const T *__restrict__ A = a + (size_t)mat * (size_t)nm * (size_t)n;
    for (int j = 0; j < n; ++j) S.m[j][k] = (k < n) ? A[k + (size_t)j * nm] : T(0);
    wsync();
    for (int c = 0; c < n; ++c) {
      const T x = (k >= c && k < n) ? S.m[c][k] : T(0);
      const T ss = group_sum<64>(x * x);
      const T xc = S.m[c][c];
      const T nrm = sqrt_(ss);
      const bool refl = (c < n - 1) && (nrm > T(0));
      const T alpha = xc >= T(0) ? -nrm : nrm;
      const T vk = (k == c) ? xc - alpha : x;
      const T tau = refl ? T(1) / (nrm * (nrm + abs_(xc))) : T(0);   // 2 / v^T v
      if (k == 0) {
        S.tau[c] = tau;
        S.sg[c] = (refl ? alpha : xc) < T(0) ? T(-1) : T(1);
      }
      S.v[k] = (k >= c && k < n) ? vk : T(0);
      wsync();
      if (refl) {
        if (k > c && k < n) {
          T s = T(0);
          for (int r = c; r < n; ++r) s = fma_(S.v[r], S.m[k][r], s);
          S.w[k] = s;
        }
        wsync();
        if (k >= c && k < n) {
          const T tv = tau * vk;
          for (int j = c + 1; j < n; ++j) S.m[j][k] = fma_(-tv, S.w[j], S.m[j][k]);
        }
      }
      // the reflector replaces column c (rows c..n-1)
      if (k >= c && k < n) S.m[c][k] = vk;
      wsync();
    }
    // Q = H_0 ... H_{n-2} I, applied backwards; reflector c touches rows and
    // columns c..n-1 only
    for (int j = 0; j < n; ++j) S.q[j][k] = (j == k) ? T(1) : T(0);
    wsync();
    for (int c = n - 2; c >= 0; --c) {
      const T tau = S.tau[c];
      if (tau == T(0)) continue;
      if (k >= c && k < n) {
        T s = T(0);
        for (int r = c; r < n; ++r) s = fma_(S.m[c][r], S.q[k][r], s);
        S.w[k] = s;
      }
      wsync();
      if (k >= c && k < n) {
        const T tv = tau * S.m[c][k];
        for (int j = c; j < n; ++j) S.q[j][k] = fma_(-tv, S.w[j], S.q[j][k]);
      }
      wsync();
    }
    T *__restrict__ Q = qout + (size_t)mat * (size_t)sq;
    if (k < n)
      for (int j = 0; j < n; ++j) Q[k + (size_t)j * ldq] = S.q[j][k] * S.sg[j];
    wsync();
  }
}

// ---- per-matrix acceptance, then the store -------------------------------------
// status[b] = ST_OVERFLOW where the Jacobi sweep cap was hit (info < 0), else 0;
// accepted matrices get V (ld n) copied into a (ld nm, first m columns).
template <typename T>
__global__ __launch_bounds__(256) void eig_warm_store_kernel(int L, int nm, int n, int m, const T *__restrict__ V,
                                                             const int *__restrict__ jinfo, int *__restrict__ status,
                                                             T *__restrict__ a, int *__restrict__ info)
{
  for (long long b = blockIdx.x; b < L; b += gridDim.x) {
    const bool bad = jinfo[b] != 0;
    if (threadIdx.x == 0) {
      status[b] = bad ? ST_OVERFLOW : 0;
      if (info && !bad) info[b] = 0;
    }
    if (bad) continue;
    const T *__restrict__ Vb = V + (size_t)b * n * n;
    T *__restrict__ Ab = a + (size_t)b * nm * n;
    for (int e = threadIdx.x; e < m * n; e += 256) {
      const int r = e % n, c = e / n;
      Ab[r + (size_t)c * nm] = Vb[e];
    }
  }
}

} // namespace eigmi
