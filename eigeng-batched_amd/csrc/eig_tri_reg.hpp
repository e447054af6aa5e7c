// eig_tri_reg.hpp - workgroup-resident Householder tridiagonalisation for
// orders up to 256: the matrix lives in the registers of one workgroup from
// the load to the last column, so HBM sees the input once and the reflectors,
// d, e and beta once (the reference's hhsy2tr keeps its trailing update in
// 3 x 3 register tiles of a thread block, hhsy2tr.hpp:197-291; here the whole
// lower triangle is register-resident).
//
// Layout: the lower triangle in 32 x 32 tiles (I, J), J <= I, numbered
// t = I (I + 1) / 2 + J; tile t belongs to wave t mod NW, slot t / NW.  In a
// tile, lane l holds row 32 I + (l & 31), columns 32 J + 16 (l >> 5) + 0..15
// (16 registers).  Diagonal tiles hold the full symmetric 32 x 32 block, so
// their row sums are complete and they contribute no column sums.
//
// Column c (c = n-1 down to 1), A = the matrix after the update of column
// c + 1, A11 = A(0:c-1, 0:c-1), in the unblocked form's convention (anorm,
// beta, u_c in W(0:c-1, c), d, e of the split kernels):
//   x = A(c, 0:c-1), y = A(c-1, 0:c-1)
//   anorm = -sign(||x||, x_{c-1}), ul = x_{c-1} - anorm, beta = 1 / (ul anorm)
//   u = x with u_{c-1} = ul
//   p = beta A11 u = beta (A11 x - anorm y)      (A11 e_{c-1} = y)
//   alpha = beta / 2 (u . p)
//   A11 += u p^T + w u^T,  w = p + 2 alpha u    (= u v^T + v u^T, v = p + alpha u)
// Two workgroup barriers per column:
//   1. (every wave) norm, per-tile row and column sums of A11 x    | barrier
//   2. (thread i) p_i from the tile sums in a fixed order, u_i     | barrier
//   3. (every wave) the rank-2 update of its tiles; the lanes of row c-2
//      publish it (the next column's y); every wave forms its own copy of the
//      next column's x = row c-1 after this update from the published
//      pre-update row and the rank-2 terms: x'_j = y_j + u_{c-1} p_j +
//      w_{c-1} u_j, so the next column starts without another barrier.
// Row sums run in the lane and across the two half-waves, column sums by a
// butterfly reduce-scatter over the 32 rows, both through lane swaps and DPP
// (no LDS traffic); the per-element sums over tiles run in a fixed order, so
// results do not depend on scheduling.
#pragma once
#include "common.hpp"

namespace eigmi {

constexpr int TR_NTB = 8;                                   // 32-row blocks: n <= 256
constexpr int TR_NTILE = TR_NTB * (TR_NTB + 1) / 2;
constexpr int TR_N = TR_NTB * 32;

template <typename T, int NW>
struct alignas(16) TriRegLds {
  alignas(16) T u[TR_N];                                    // u of the current column, zero beyond c-1
  alignas(16) T p[TR_N];
  alignas(16) T y[2][TR_N];                                 // y of column c in buffer c & 1
  alignas(16) T xw[NW][TR_N];                               // each wave's copy of x, zero beyond c-1
  alignas(16) T prow[TR_NTILE][32];                         // per-tile row sums
  alignas(16) T pcol[TR_NTILE][32];                         // per-tile column sums
  alignas(16) T d[TR_N];
  alignas(16) T e[TR_N];
  alignas(16) T beta[TR_N];
  alignas(16) T red[TR_NTB / 2];
};

// The lanes holding row r write y[j] = A(r, j), j <= r, into yb.  Only the
// tiles of block row r / 32 take part (a wave-uniform test); off-diagonal
// tiles store their 16 values whole, the diagonal one compares against r
// itself, so no per-lane mask is hoisted out of the column loop.
template <typename T, int KT>
__device__ __forceinline__ void tr_publish(int r, const T (&A)[KT][16], const int (&tI)[KT], const int (&tJ)[KT],
                                           int rr, int hh, T *__restrict__ yb)
{
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int I = tI[k], J = tJ[k];
    const int col0 = 32 * J + 16 * hh;
    if (r >= 0 && I == (r >> 5) && rr == (r & 31)) {
      if (J < I) {
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) yb[col0 + cc] = A[k][cc];
      } else {
        const int lim = r - col0;                           // col0 + cc <= r  <=>  cc <= lim
#pragma unroll
        for (int cc = 0; cc < 16; ++cc)
          if (cc <= lim) yb[col0 + cc] = A[k][cc];
      }
    }
  }
}

// Every wave: its copy of x for column c (= row c of the matrix after the
// update of column c + 1): x_j = y_j + uc p_j + wc u_j for j < c, 0 beyond;
// returns the lane's part of ||x||^2.
template <typename T, int NW>
__device__ __forceinline__ T tr_form_x(int c, T uc, T wc, const T *__restrict__ yb, T *__restrict__ xw,
                                       const TriRegLds<T, NW> &S, int lane)
{
  T sg = T(0);
#pragma unroll
  for (int m = 0; m < TR_N / 64; ++m) {
    const int j = lane + 64 * m;
    const T v = fma_(wc, S.u[j], fma_(uc, S.p[j], yb[j]));
    const T x = (j < c) ? v : T(0);
    xw[j] = x;
    sg = fma_(x, x, sg);
  }
  return sg;
}

// NW waves, KT tile slots per wave (NW KT >= tiles of the order); WPE waves
// per SIMD (the register budget: 512 / WPE VGPRs per lane).
template <typename T, int NW, int KT, int WPE>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE))) void eig_tri_reg_kernel(
    int L, int nm, int n, const T *__restrict__ a, T *__restrict__ W, T *__restrict__ de, int ld,
    T *__restrict__ bws)
{
  constexpr int NT = 64 * NW;
  __shared__ TriRegLds<T, NW> S;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int rr = lane & 31, hh = lane >> 5;
  const int nb = (n + 31) >> 5;
  const int ntile = nb * (nb + 1) / 2;
  const int nsum = (nb + 1) >> 1;                           // waves that form p
  int tI[KT], tJ[KT], tT[KT];
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int t = wv + NW * k;
    int I = 0;
    while ((I + 1) * (I + 2) / 2 <= t) ++I;
    tT[k] = t;
    tI[k] = (t < ntile) ? I : TR_NTB + 1;                   // out of range: never active
    tJ[k] = (t < ntile) ? t - I * (I + 1) / 2 : 0;
  }
  T *__restrict__ xw = S.xw[wv];
  T A[KT][16];

  for (long long mat = blockIdx.x; mat < L; mat += gridDim.x) {
    // n and nm re-read opaquely per matrix: otherwise the per-element load
    // addresses are hoisted out of the batch loop and held in registers for
    // the whole kernel
    int nv = n, nmv = nm;
    asm volatile("" : "+s"(nv), "+s"(nmv));
    const T *__restrict__ Ain = a + (size_t)mat * (size_t)nm * (size_t)n;
    T *__restrict__ Wm = W + (size_t)mat * (size_t)n * (size_t)n;
    // lower (row, col) = Ain(min, max): the input's upper triangle.  Rows and
    // columns beyond n (and unused slots) read a clamped in-range element: they
    // only ever meet zero entries of x, u and p.
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int col0 = 32 * tJ[k] + 16 * hh;
      if (tJ[k] < tI[k] && 32 * tI[k] + 32 <= nv) {
        // interior off-diagonal tile: 16 consecutive elements of input column `row`
        const T *__restrict__ src = Ain + (size_t)(32 * tI[k] + rr) * nmv + col0;
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) A[k][cc] = src[cc];
      } else {
        const int row = min(32 * tI[k] + rr, nv - 1);
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
          const int col = min(col0 + cc, nv - 1);
          const int lo = row < col ? row : col, hi = row < col ? col : row;
          A[k][cc] = Ain[lo + (size_t)hi * nmv];
        }
      }
    }
    for (int i = tid; i < TR_N; i += NT) {
      S.u[i] = T(0);
      S.p[i] = T(0);
      S.e[i] = T(0);
      S.beta[i] = T(0);
    }
    // rows n-1 (x of the first column) and n-2 (its y)
    tr_publish<T, KT>(nv - 1, A, tI, tJ, rr, hh, S.y[nv & 1]);
    tr_publish<T, KT>(nv - 2, A, tI, tJ, rr, hh, S.y[(nv - 1) & 1]);
    __syncthreads();
    T sgl = tr_form_x<T, NW>(nv - 1, T(0), T(0), S.y[nv & 1], xw, S, lane);
    if (tid == 0) S.d[nv - 1] = S.y[nv & 1][nv - 1];

    for (int c = nv - 1; c >= 1; --c) {
      // the lane-derived indices are formed again for every column: held from
      // before the loop they were spilled, and their reloads from scratch sat
      // on the column's critical path
      int lane_c = lane;
      asm volatile("" : "+v"(lane_c));
      const int rr = lane_c & 31, hh = lane_c >> 5, tid_c = wv * 64 + lane_c;
      // 1. norm, per-tile row and column sums of A11 x (x in this wave's copy)
      wsync();
      const T sg = group_sum<64>(sgl);
      const T xl = xw[c - 1];
      const T anorm = -sign_(sqrt_(sg), xl);
      const T ul = xl - anorm;
      const T beta = T(1) / nz(ul * anorm);
      const int ib = (c - 1) >> 5;                          // last active block
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        if (tI[k] <= ib) {
          const int I = tI[k], J = tJ[k], t = tT[k];
          const int row = 32 * I + rr, col0 = 32 * J + 16 * hh;
          T s0 = T(0), s1 = T(0);
#pragma unroll
          for (int cc = 0; cc < 16; cc += 2) {
            s0 = fma_(A[k][cc], xw[col0 + cc], s0);
            s1 = fma_(A[k][cc + 1], xw[col0 + cc + 1], s1);
          }
          const T s = swap_sum<32>(s0 + s1, s0 + s1);       // both halves of the row
          if (hh == 0) S.prow[t][rr] = s;
          if (I != J) {
            const T ur = xw[row];
            T tc[16];
#pragma unroll
            for (int cc = 0; cc < 16; ++cc) tc[cc] = A[k][cc] * ur;
            // column sums over the 32 rows: reduce-scatter, lane rr ends
            // with column rr >> 1 of its half
#pragma unroll
            for (int i = 0; i < 8; ++i) tc[i] = swap_sum<16>(tc[i], tc[i + 8]);
            reduce_scatter_step<4>(tc, (rr & 8) != 0, MvRM());
            reduce_scatter_step<2>(tc, (rr & 4) != 0, MvHM());
            reduce_scatter_step<1>(tc, (rr & 2) != 0, MvX2());
            tc[0] += lane_move(tc[0], MvX1());
            if (!(rr & 1)) S.pcol[t][16 * hh + (rr >> 1)] = tc[0];
          }
        }
      }
      __syncthreads();                                      // tile sums complete
      // 2. p_i (thread i), u_i, the partial sums of u . p
      if (wv < nsum) {
        const int i = tid_c;
        const int b = i >> 5, r = i & 31;
        T s = T(0);
#pragma unroll
        for (int J = 0; J < TR_NTB; ++J) {
          const T v = S.prow[b * (b + 1) / 2 + J][r];      // in range for every J < 8
          s += (J <= b) ? v : T(0);
        }
#pragma unroll
        for (int I = 1; I < TR_NTB; ++I) {
          const T v = S.pcol[I * (I + 1) / 2 + b][r];
          s += (I > b && I <= ib) ? v : T(0);
        }
        const T x = xw[i];
        const T pi = (i < c) ? beta * fma_(-anorm, S.y[c & 1][i], s) : T(0);
        const T ui = (i == c - 1) ? ul : x;                 // x is 0 beyond c-1
        S.p[i] = pi;
        S.u[i] = ui;
        const T pr = group_sum<64>(pi * ui);
        if (lane_c == 0) S.red[wv] = pr;
      }
      if (tid_c == 0) {
        S.e[c - 1] = anorm;
        S.beta[c] = beta;
      }
      __syncthreads();                                      // p, u complete
      T alpha = T(0);
      for (int q = 0; q < nsum; ++q) alpha += S.red[q];
      alpha *= beta * T(0.5);
      const T alpha2 = alpha + alpha;
      if (wv == c % NW)
        for (int r = lane_c; r < c; r += 64) Wm[r + (size_t)c * n] = S.u[r];
      // 3. rank-2 update of the active tiles
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        if (tI[k] <= ib) {
          const int row = 32 * tI[k] + rr, col0 = 32 * tJ[k] + 16 * hh;
          const T ui = S.u[row];
          const T wi = fma_(alpha2, ui, S.p[row]);
#pragma unroll
          for (int cc = 0; cc < 16; ++cc)
            A[k][cc] = fma_(ui, S.p[col0 + cc], fma_(wi, S.u[col0 + cc], A[k][cc]));
        }
      }
      // row c-2 after this update: y of column c-1
      tr_publish<T, KT>(c - 2, A, tI, tJ, rr, hh, S.y[(c - 1) & 1]);
      // x of column c-1 = row c-1 after this update, from its pre-update y;
      // d_{c-1} likewise
      const T uc = S.u[c - 1];
      const T wc = fma_(alpha2, uc, S.p[c - 1]);
      sgl = tr_form_x<T, NW>(c - 1, uc, wc, S.y[c & 1], xw, S, lane_c);
      if (tid_c == 0) S.d[c - 1] = fma_(wc, uc, fma_(uc, S.p[c - 1], S.y[c & 1][c - 1]));
    }
    __syncthreads();
    for (int i = tid; i < nv; i += NT) {
      de[(size_t)i * ld + mat] = S.d[i];
      de[(size_t)(nv + i) * ld + mat] = S.e[i];
      bws[(size_t)mat * nv + i] = S.beta[i];
    }
    __syncthreads();
  }
}

} // namespace eigmi
