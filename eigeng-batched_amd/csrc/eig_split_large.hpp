// eig_split_large.hpp - split solver for the orders beyond the register kernels
// (FP32 128 < n <= 512, FP64 64 < n <= 512).
//
//   eig_tri_large_kernel   (workgroup per matrix, persistent grid): copy a into
//                          a workspace slot, Householder tridiagonalisation there
//                          (u_c in W(0:c-1, c), beta_c to bws), d / e to the
//                          matrix-minor layout read by the QL kernel; a itself
//                          is not written, so the workgroup fallback can still
//                          solve any matrix whose rotation log overflows.
//   eig_ql_lane_kernel     (eig_split.hpp, NLANE matrices per wave): the QL
//                          recurrence with rotation and sweep logs, sort, w.
//   eig_apply_large_kernel (workgroup per block of 64 RPL rows of a matrix):
//                          rows of Q = H_{n-1} ... H_1 formed in registers, then
//                          every logged sweep replayed on them, then the
//                          permuted store.
//
// Apply kernel layout: lane = row of the block (RPL rows per lane), wave q = CW-column chunk
// [CW q, CW q + CW-1] held in registers (the description below uses CW = 32).  Rotations act on columns, so rows are
// independent, but one sweep chains through the columns from the top of its
// window down: wave q handles the pairs i in [32q-1, 32q+30] (pair i finalises
// column i+1, i.e. the wave's own columns).  Pair 32q-1 needs the top column of
// wave q-1 (published in LDS after that wave's previous sweep) and hands the
// pending value down to wave q-1 (LDS carry).  With step T(q, s) = 2 s +
// (NCH-1-q) for sweep s, wave q-1 has finished sweep s-1 one step before wave
// q starts sweep s, and starts sweep s one step after it: one workgroup barrier
// per step keeps the chain exact.  The rotation of a pair and its arithmetic are
// those of apply_sweep_rotations (the n <= 64 apply kernel).
#pragma once
#include "common.hpp"
#include "eig_global.hpp"
#include "eig_split.hpp"
#include <cstddef>

namespace eigmi {

// ---- blocked Householder tridiagonalisation ---------------------------------
// The reduction of column c needs A11 = A(0:c-1, 0:c-1) with the rank-2
// updates of all earlier columns applied.  In panels of NBP columns those
// updates are kept as the vectors U_k, V_k (zero beyond their length) and
// folded in analytically (the dlatrd arrangement):
//   column c      = A(:, c) + sum_k (U_k V_k(c) + V_k U_k(c))
//   p = beta A11 u = beta (A u + sum_k (U_k (V_k . u) + V_k (U_k . u)))
// and A11 itself is updated once per panel.  Per column this is one read pass
// over A11 instead of a read pass plus a read-modify-write pass; the
// Householder convention (anorm, beta, u_c stored in A(0:c-1, c), d, e) is the
// unblocked one's.
template <typename T, int NBP, int NMAXT, int RG>
struct alignas(16) TriBlkLds {
  alignas(16) T U[NBP][NMAXT];
  alignas(16) T V[NBP][NMAXT];
  alignas(16) T u[NMAXT];
  alignas(16) T pp[RG][NMAXT];                   // per-column-group partial products
  alignas(16) T red[16][2 * NBP + 2];
  alignas(16) T d[NMAXT];
  alignas(16) T e[NMAXT];
  alignas(16) T beta[NMAXT];
};

// Sum over the workgroup of K per-thread values (K <= 2 NBP + 2); every thread
// gets the K sums.
template <int NT, int K, typename S_t, typename T>
__device__ __forceinline__ void block_sums(T (&x)[K], int cnt, S_t &S)
{
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    if (i < cnt) {
      const T v = group_sum<64>(x[i]);
      if (lane == 0) S.red[wv][i] = v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < K; ++i) {
    if (i < cnt) {
      T sum = T(0);
#pragma unroll
      for (int q = 0; q < NT / 64; ++q) sum += S.red[q][i];
      x[i] = sum;
    }
  }
  __syncthreads();
}

// Thread t works on rows t % RT and t % RT + RT (RT = NT / RG) and, in the two
// passes over the trailing matrix, on the columns j = g (mod RG), g = t / RT.
template <typename T, int NT, int NBP, int NMAXT, int RG>
__device__ void blocked_tridiag(T *__restrict__ A, int nm, int n, TriBlkLds<T, NBP, NMAXT, RG> &S)
{
  constexpr int RT = NT / RG;
  static_assert(2 * RT >= NMAXT, "two rows per thread cover the matrix");
  const int t = threadIdx.x;
  const int g = t / RT;
  const int r0 = t - g * RT, r1 = r0 + RT;
  const bool lead = (g == 0);                        // the group that owns per-row state
  // symmetrise (lower <- upper)
  for (int c = 0; c < n; ++c)
    for (int r = c + 1 + t; r < n; r += NT) A[r + (size_t)c * nm] = A[c + (size_t)r * nm];
  for (int i = t; i < NMAXT; i += NT) {
#pragma unroll
    for (int k = 0; k < NBP; ++k) { S.U[k][i] = T(0); S.V[k][i] = T(0); }
  }
  __syncthreads();

  for (int ctop = n - 1; ctop >= 1; ctop -= NBP) {
    const int nbk = (NBP < ctop) ? NBP : ctop;
    for (int k = 0; k < nbk; ++k) {
      const int c = ctop - k;
      // 1. column c of the updated matrix (rows < c) and its diagonal
      T uc0 = T(0), uc1 = T(0);
      if (lead) {
        T ukc[NBP], vkc[NBP];
#pragma unroll
        for (int kk = 0; kk < NBP; ++kk) { ukc[kk] = S.U[kk][c]; vkc[kk] = S.V[kk][c]; }
        if (r0 <= c) {
          T acc = A[r0 + (size_t)c * nm];
#pragma unroll
          for (int kk = 0; kk < NBP; ++kk)
            if (kk < k) acc = fma_(S.U[kk][r0], vkc[kk], fma_(S.V[kk][r0], ukc[kk], acc));
          if (r0 < c) uc0 = acc; else S.d[c] = acc;
        }
        if (r1 <= c) {
          T acc = A[r1 + (size_t)c * nm];
#pragma unroll
          for (int kk = 0; kk < NBP; ++kk)
            if (kk < k) acc = fma_(S.U[kk][r1], vkc[kk], fma_(S.V[kk][r1], ukc[kk], acc));
          if (r1 < c) uc1 = acc; else S.d[c] = acc;
        }
        if (r0 < c) S.u[r0] = uc0;
        if (r1 < c) S.u[r1] = uc1;
      }
      T sg[1] = {uc0 * uc0 + uc1 * uc1};
      block_sums<NT, 1>(sg, 1, S);
      const T ulast = S.u[c - 1];
      const T anorm = -sign_(sqrt_(sg[0]), ulast);
      const T ul = ulast - anorm;
      const T beta = T(1) / nz(ul * anorm);
      if (r0 == c - 1) uc0 = ul;
      if (r1 == c - 1) uc1 = ul;
      // u_c is stored in A(0:c-1, c) (with u(c-1) = ul), as in the unblocked form
      if (lead) {
        if (r0 < c) A[r0 + (size_t)c * nm] = uc0;
        if (r1 < c) A[r1 + (size_t)c * nm] = uc1;
      }
      __syncthreads();
      if (t == 0) {
        S.u[c - 1] = ul;
        S.e[c - 1] = anorm;
        S.beta[c] = beta;
      }
      __syncthreads();
      // 2. dots of the panel vectors with u, then p
      T dd[2 * NBP];
#pragma unroll
      for (int kk = 0; kk < 2 * NBP; ++kk) dd[kk] = T(0);
      if (lead) {
#pragma unroll
        for (int kk = 0; kk < NBP; ++kk) {
          if (kk < k) {
            if (r0 < c) { dd[2 * kk] = S.V[kk][r0] * uc0; dd[2 * kk + 1] = S.U[kk][r0] * uc0; }
            if (r1 < c) {
              dd[2 * kk] = fma_(S.V[kk][r1], uc1, dd[2 * kk]);
              dd[2 * kk + 1] = fma_(S.U[kk][r1], uc1, dd[2 * kk + 1]);
            }
          }
        }
      }
      block_sums<NT, 2 * NBP>(dd, 2 * k, S);
      // A(0:c-1, 0:c-1) u: column group g sums over j = g (mod RG)
      T p0 = T(0), p1 = T(0);
      if (r0 < c) {
        for (int j = g; j < c; j += RG) {
          const T uj = S.u[j];
          p0 = fma_(A[r0 + (size_t)j * nm], uj, p0);
          if (r1 < c) p1 = fma_(A[r1 + (size_t)j * nm], uj, p1);
        }
      }
      if (RG > 1) {
        S.pp[g][r0] = p0;
        if (r1 < NMAXT) S.pp[g][r1] = p1;
        __syncthreads();
        p0 = T(0);
        p1 = T(0);
        if (lead) {
#pragma unroll
          for (int gg = 0; gg < RG; ++gg) {
            p0 += S.pp[gg][r0];
            if (r1 < NMAXT) p1 += S.pp[gg][r1];
          }
        }
      }
      if (lead) {
#pragma unroll
        for (int kk = 0; kk < NBP; ++kk) {
          if (kk < k) {
            if (r0 < c) p0 = fma_(S.U[kk][r0], dd[2 * kk], fma_(S.V[kk][r0], dd[2 * kk + 1], p0));
            if (r1 < c) p1 = fma_(S.U[kk][r1], dd[2 * kk], fma_(S.V[kk][r1], dd[2 * kk + 1], p1));
          }
        }
      }
      p0 = (lead && r0 < c) ? p0 * beta : T(0);
      p1 = (lead && r1 < c) ? p1 * beta : T(0);
      T al[1] = {p0 * uc0 + p1 * uc1};
      block_sums<NT, 1>(al, 1, S);
      const T alpha = al[0] * (beta * T(0.5));
      if (lead) {
        if (r0 < NMAXT) { S.U[k][r0] = (r0 < c) ? uc0 : T(0); S.V[k][r0] = (r0 < c) ? fma_(alpha, uc0, p0) : T(0); }
        if (r1 < NMAXT) { S.U[k][r1] = (r1 < c) ? uc1 : T(0); S.V[k][r1] = (r1 < c) ? fma_(alpha, uc1, p1) : T(0); }
      }
      __syncthreads();
    }
    // 3. panel update of the trailing block A(0:clow-1, 0:clow-1)
    const int clow = ctop - nbk + 1;
    {
      T ur0[NBP], vr0[NBP], ur1[NBP], vr1[NBP];
#pragma unroll
      for (int kk = 0; kk < NBP; ++kk) {
        ur0[kk] = (r0 < clow && kk < nbk) ? S.U[kk][r0] : T(0);
        vr0[kk] = (r0 < clow && kk < nbk) ? S.V[kk][r0] : T(0);
        ur1[kk] = (r1 < clow && kk < nbk) ? S.U[kk][r1] : T(0);
        vr1[kk] = (r1 < clow && kk < nbk) ? S.V[kk][r1] : T(0);
      }
      if (r0 < clow) {
        for (int j = g; j < clow; j += RG) {
          T a0 = A[r0 + (size_t)j * nm];
          T a1 = (r1 < clow) ? A[r1 + (size_t)j * nm] : T(0);
#pragma unroll
          for (int kk = 0; kk < NBP; ++kk) {
            if (kk < nbk) {
              const T uj = S.U[kk][j], vj = S.V[kk][j];
              a0 = fma_(ur0[kk], vj, fma_(vr0[kk], uj, a0));
              a1 = fma_(ur1[kk], vj, fma_(vr1[kk], uj, a1));
            }
          }
          A[r0 + (size_t)j * nm] = a0;
          if (r1 < clow) A[r1 + (size_t)j * nm] = a1;
        }
      }
    }
    __syncthreads();
    for (int i = t; i < NMAXT; i += NT) {
#pragma unroll
      for (int k = 0; k < NBP; ++k) { S.U[k][i] = T(0); S.V[k][i] = T(0); }
    }
    __syncthreads();
  }
  if (t == 0) { S.d[0] = A[0]; S.e[n - 1] = T(0); }
  __syncthreads();
}

// Symmetric-storage form of blocked_tridiag (the default for n <= 256): the
// live matrix is the lower triangle with the diagonal, loaded as the transpose
// of the input's upper triangle.  The SYMV reads each lower element once per
// column (32 x 32 tiles, row and column sums) and the panel update touches only
// the lower triangle, so both passes move half the bytes of the full form.
// The strictly upper part of A receives the u_c storage and nothing else; the
// Householder arithmetic and outputs are those of blocked_tridiag.
template <typename T, int NT, int NBP, int NMAXT, int RG>
__device__ void blocked_tridiag_sym(const T *__restrict__ Ain, int nma, T *__restrict__ A, int nm, int n,
                                    TriBlkLds<T, NBP, NMAXT, RG> &S, T (&pw)[NT / 64][NMAXT])
{
  constexpr int RT = NT / RG;
  static_assert(2 * RT >= NMAXT, "two rows per thread cover the matrix");
  // the transpose staging uses U and V, which are adjacent
  static_assert(NT % 32 == 0 && 2 * NBP * NMAXT >= 32 * 33, "transpose staging fits in U and V");
  typedef TriBlkLds<T, NBP, NMAXT, RG> Lds_t;
  static_assert(offsetof(Lds_t, V) == offsetof(Lds_t, U) + sizeof(T) * NBP * NMAXT, "V follows U");
  const int t = threadIdx.x;
  const int g = t / RT;
  const int r0 = t - g * RT, r1 = r0 + RT;
  const bool lead = (g == 0);                        // the group that owns per-row state
  // lower triangle of A <- transpose of the upper triangle of Ain, in 32 x 32
  // tiles staged through LDS (coalesced on both sides)
  {
    T (*stg)[33] = reinterpret_cast<T (*)[33]>(&S.U[0][0]);
    const int nt = (n + 31) >> 5, tr = t & 31;
    for (int I = 0; I < nt; ++I) {
      for (int J = 0; J <= I; ++J) {
        for (int tc = t >> 5; tc < 32; tc += NT / 32) {
          const int ar = 32 * J + tr, ac = 32 * I + tc;   // Ain(ar, ac), upper block (J, I)
          if (ar < n && ac < n && (I > J || tr <= tc)) stg[tr][tc] = Ain[ar + (size_t)ac * nma];
        }
        __syncthreads();
        for (int tc = t >> 5; tc < 32; tc += NT / 32) {
          const int wr = 32 * I + tr, wc = 32 * J + tc;   // A(wr, wc) = Ain(wc, wr)
          if (wr < n && wc < n && (I > J || tr >= tc)) A[wr + (size_t)wc * nm] = stg[tc][tr];
        }
        __syncthreads();
      }
    }
  }
  for (int i = t; i < NMAXT; i += NT) {
#pragma unroll
    for (int k = 0; k < NBP; ++k) { S.U[k][i] = T(0); S.V[k][i] = T(0); }
  }
  __syncthreads();

  for (int ctop = n - 1; ctop >= 1; ctop -= NBP) {
    const int nbk = (NBP < ctop) ? NBP : ctop;
    for (int k = 0; k < nbk; ++k) {
      const int c = ctop - k;
      // 1. column c of the updated matrix (rows < c) and its diagonal
      T uc0 = T(0), uc1 = T(0);
      if (lead) {
        T ukc[NBP], vkc[NBP];
#pragma unroll
        for (int kk = 0; kk < NBP; ++kk) { ukc[kk] = S.U[kk][c]; vkc[kk] = S.V[kk][c]; }
        if (r0 <= c) {
          T acc = A[c + (size_t)r0 * nm];              // row c of the lower triangle
#pragma unroll
          for (int kk = 0; kk < NBP; ++kk)
            if (kk < k) acc = fma_(S.U[kk][r0], vkc[kk], fma_(S.V[kk][r0], ukc[kk], acc));
          if (r0 < c) uc0 = acc; else S.d[c] = acc;
        }
        if (r1 <= c) {
          T acc = A[c + (size_t)r1 * nm];
#pragma unroll
          for (int kk = 0; kk < NBP; ++kk)
            if (kk < k) acc = fma_(S.U[kk][r1], vkc[kk], fma_(S.V[kk][r1], ukc[kk], acc));
          if (r1 < c) uc1 = acc; else S.d[c] = acc;
        }
        if (r0 < c) S.u[r0] = uc0;
        if (r1 < c) S.u[r1] = uc1;
      }
      T sg[1] = {uc0 * uc0 + uc1 * uc1};
      block_sums<NT, 1>(sg, 1, S);
      const T ulast = S.u[c - 1];
      const T anorm = -sign_(sqrt_(sg[0]), ulast);
      const T ul = ulast - anorm;
      const T beta = T(1) / nz(ul * anorm);
      if (r0 == c - 1) uc0 = ul;
      if (r1 == c - 1) uc1 = ul;
      // u_c is stored in A(0:c-1, c) (with u(c-1) = ul), as in the unblocked form
      if (lead) {
        if (r0 < c) A[r0 + (size_t)c * nm] = uc0;
        if (r1 < c) A[r1 + (size_t)c * nm] = uc1;
      }
      __syncthreads();
      if (t == 0) {
        S.u[c - 1] = ul;
        S.e[c - 1] = anorm;
        S.beta[c] = beta;
      }
      __syncthreads();
      // 2. p = beta (A11 u + panel corrections)
      // A(0:c-1, 0:c-1) u from the lower triangle only, in 32x32 tiles (I, J),
      // J <= I: a tile adds T u_J to p_I (row sums) and, off the diagonal,
      // T^T u_I to p_J (column sums over the tile's rows).  Each wave keeps its
      // own partial vector pw[wave], summed in wave order afterwards, so the
      // result does not depend on scheduling.
      {
        const int lane = t & 63, wv = t >> 6, rr = lane & 31, hh = lane >> 5;
        for (int i = lane; i < c; i += 64) pw[wv][i] = T(0);
        wsync();
        const int nt = (c + 31) >> 5;
        const int ntiles = nt * (nt + 1) / 2;
        for (int tile = wv; tile < ntiles; tile += NT / 64) {
          int I = 0;
          while ((I + 1) * (I + 2) / 2 <= tile) ++I;
          const int J = tile - I * (I + 1) / 2;
          const int row = 32 * I + rr;
          const int col0 = 32 * J + 16 * hh;
          const T ur = (row < c) ? S.u[row] : T(0);
          // every load is issued (clamped into the live block, value dropped
          // when not live) so the 16 loads go out back to back
          const T *__restrict__ arow = A + (row < c ? row : c - 1);
          T srow = T(0);
          T tc[16];
          T av[16];
#pragma unroll
          for (int cc = 0; cc < 16; ++cc) {
            const int col = col0 + cc;
            av[cc] = arow[(size_t)(col < c ? col : c - 1) * nm];
          }
#pragma unroll
          for (int cc = 0; cc < 16; ++cc) {
            const int col = col0 + cc;
            const bool live = row < c && col < c && (I > J || col <= row);
            const T v = live ? av[cc] : T(0);
            const T uc = (col < c) ? S.u[col] : T(0);
            srow = fma_(v, uc, srow);
            tc[cc] = (I > J || col < row) ? v * ur : T(0);
          }
          srow += lane_move(srow, MvX32());
          if (hh == 0 && row < c) pw[wv][row] += srow;
          wsync();
          // column sums over the 32 rows: reduce-scatter, lane rr ends with
          // column rr >> 1 of its half
          reduce_scatter_step<8>(tc, (rr & 16) != 0, MvX16());
          reduce_scatter_step<4>(tc, (rr & 8) != 0, MvRM());
          reduce_scatter_step<2>(tc, (rr & 4) != 0, MvHM());
          reduce_scatter_step<1>(tc, (rr & 2) != 0, MvX2());
          tc[0] += lane_move(tc[0], MvX1());
          if (!(rr & 1) && col0 + (rr >> 1) < c) pw[wv][col0 + (rr >> 1)] += tc[0];
          wsync();
        }
      }
      __syncthreads();
      // dots of the panel vectors with u
      T dd[2 * NBP];
#pragma unroll
      for (int kk = 0; kk < 2 * NBP; ++kk) dd[kk] = T(0);
      if (lead) {
#pragma unroll
        for (int kk = 0; kk < NBP; ++kk) {
          if (kk < k) {
            if (r0 < c) { dd[2 * kk] = S.V[kk][r0] * uc0; dd[2 * kk + 1] = S.U[kk][r0] * uc0; }
            if (r1 < c) {
              dd[2 * kk] = fma_(S.V[kk][r1], uc1, dd[2 * kk]);
              dd[2 * kk + 1] = fma_(S.U[kk][r1], uc1, dd[2 * kk + 1]);
            }
          }
        }
      }
      block_sums<NT, 2 * NBP>(dd, 2 * k, S);
      T p0 = T(0), p1 = T(0);
      if (lead) {
#pragma unroll
        for (int q = 0; q < NT / 64; ++q) {
          if (r0 < c) p0 += pw[q][r0];
          if (r1 < c) p1 += pw[q][r1];
        }
      }
      if (lead) {
#pragma unroll
        for (int kk = 0; kk < NBP; ++kk) {
          if (kk < k) {
            if (r0 < c) p0 = fma_(S.U[kk][r0], dd[2 * kk], fma_(S.V[kk][r0], dd[2 * kk + 1], p0));
            if (r1 < c) p1 = fma_(S.U[kk][r1], dd[2 * kk], fma_(S.V[kk][r1], dd[2 * kk + 1], p1));
          }
        }
      }
      p0 = (lead && r0 < c) ? p0 * beta : T(0);
      p1 = (lead && r1 < c) ? p1 * beta : T(0);
      T al[1] = {p0 * uc0 + p1 * uc1};
      block_sums<NT, 1>(al, 1, S);
      const T alpha = al[0] * (beta * T(0.5));
      if (lead) {
        if (r0 < NMAXT) { S.U[k][r0] = (r0 < c) ? uc0 : T(0); S.V[k][r0] = (r0 < c) ? fma_(alpha, uc0, p0) : T(0); }
        if (r1 < NMAXT) { S.U[k][r1] = (r1 < c) ? uc1 : T(0); S.V[k][r1] = (r1 < c) ? fma_(alpha, uc1, p1) : T(0); }
      }
      __syncthreads();
    }
    // 3. panel update of the trailing block A(0:clow-1, 0:clow-1)
    const int clow = ctop - nbk + 1;
    {
      T ur0[NBP], vr0[NBP], ur1[NBP], vr1[NBP];
#pragma unroll
      for (int kk = 0; kk < NBP; ++kk) {
        ur0[kk] = (r0 < clow && kk < nbk) ? S.U[kk][r0] : T(0);
        vr0[kk] = (r0 < clow && kk < nbk) ? S.V[kk][r0] : T(0);
        ur1[kk] = (r1 < clow && kk < nbk) ? S.U[kk][r1] : T(0);
        vr1[kk] = (r1 < clow && kk < nbk) ? S.V[kk][r1] : T(0);
      }
      if (r0 < clow) {
        const int jend = ((r1 < clow) ? r1 : r0) + 1;  // lower triangle and diagonal only
        for (int j = g; j < jend; j += RG) {
          const bool w0 = j <= r0;
          T a0 = w0 ? A[r0 + (size_t)j * nm] : T(0);
          T a1 = (r1 < clow) ? A[r1 + (size_t)j * nm] : T(0);
#pragma unroll
          for (int kk = 0; kk < NBP; ++kk) {
            if (kk < nbk) {
              const T uj = S.U[kk][j], vj = S.V[kk][j];
              a0 = fma_(ur0[kk], vj, fma_(vr0[kk], uj, a0));
              a1 = fma_(ur1[kk], vj, fma_(vr1[kk], uj, a1));
            }
          }
          if (w0) A[r0 + (size_t)j * nm] = a0;
          if (r1 < clow) A[r1 + (size_t)j * nm] = a1;
        }
      }
    }
    __syncthreads();
    for (int i = t; i < NMAXT; i += NT) {
#pragma unroll
      for (int k = 0; k < NBP; ++k) { S.U[k][i] = T(0); S.V[k][i] = T(0); }
    }
    __syncthreads();
  }
  if (t == 0) { S.d[0] = A[0]; S.e[n - 1] = T(0); }
  __syncthreads();
}

template <typename T, int NT, int NBP, int NMAXT, int RG, bool SYM>
__device__ __forceinline__ void tri_large_one(const T *__restrict__ A, int nm, int n, T *__restrict__ Wm,
                                              TriBlkLds<T, NBP, NMAXT, RG> &S,
                                              T (&pw)[SYM ? NT / 64 : 1][SYM ? NMAXT : 1])
{
  const int t = threadIdx.x;
  if constexpr (SYM) {
    blocked_tridiag_sym<T, NT, NBP, NMAXT, RG>(A, nm, Wm, n, n, S, pw);
  } else {
    for (int c = 0; c < n; ++c)
      for (int r = t; r <= c; r += NT) Wm[r + (size_t)c * n] = A[r + (size_t)c * nm];
    __syncthreads();
    blocked_tridiag<T, NT, NBP, NMAXT, RG>(Wm, n, n, S);
  }
}

// occupancy floor of the tridiagonalisation (waves per SIMD): the FP64
// lower-triangle form is held to 128 VGPRs (four 4-wave workgroups per CU)
template <typename T, bool SYM> struct TriWaves { static constexpr int value = (SYM && sizeof(T) == 8) ? 4 : 1; };
// panel width of the lower-triangle form: 4 for FP64 (the LDS of four
// workgroups per CU), 8 for FP32
template <typename T> struct TriNbp { static constexpr int value = sizeof(T) == 8 ? 4 : 8; };

template <typename T, int NT, int NBP, int NMAXT, int RG, bool SYM = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(TriWaves<T, SYM>::value))) void eig_tri_large_kernel(int L, int nm, int n, const T *__restrict__ a,
                                                           T *__restrict__ W, T *__restrict__ de, int ld,
                                                           T *__restrict__ bws)
{
  __shared__ TriBlkLds<T, NBP, NMAXT, RG> S;
  __shared__ T pw[SYM ? NT / 64 : 1][SYM ? NMAXT : 1];
  const int t = threadIdx.x;
  for (long long mat = blockIdx.x; mat < L; mat += gridDim.x) {
    const T *__restrict__ A = a + (size_t)mat * (size_t)nm * (size_t)n;
    T *__restrict__ Wm = W + (size_t)mat * (size_t)n * (size_t)n;
    tri_large_one<T, NT, NBP, NMAXT, RG, SYM>(A, nm, n, Wm, S, pw);
    for (int i = t; i < n; i += NT) {
      de[(size_t)i * ld + mat] = S.d[i];
      de[(size_t)(n + i) * ld + mat] = S.e[i];
      bws[(size_t)mat * n + i] = S.beta[i];
    }
    __syncthreads();
  }
}

// Uneven-batch form: slot idx holds list entry idx of the chunk (order n from
// nv, packed input at a + aoff, W slot nmx^2, d / e at rows i / nmx + i of de).
template <typename T, int NT, int NBP, int NMAXT, int RG, bool SYM = true>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(TriWaves<T, SYM>::value))) void eig_tri_large_list_kernel(UvSplitArgs u, const T *__restrict__ a,
                                                                T *__restrict__ W, T *__restrict__ de,
                                                                T *__restrict__ bws)
{
  __shared__ TriBlkLds<T, NBP, NMAXT, RG> S;
  __shared__ T pw[SYM ? NT / 64 : 1][SYM ? NMAXT : 1];
  const int t = threadIdx.x;
  const int ld = u.cap, nmx = u.nmx;
  for (long long idx = blockIdx.x; idx < u.cap; idx += gridDim.x) {
    int mat;
    if (!uv_split_entry(u, idx, mat)) continue;       // the same for the whole workgroup
    const int n = __builtin_amdgcn_readfirstlane(u.nv[mat]);
    T *__restrict__ Wm = W + (size_t)idx * nmx * nmx;
    tri_large_one<T, NT, NBP, NMAXT, RG, SYM>(a + u.aoff[mat], n, n, Wm, S, pw);
    for (int i = t; i < n; i += NT) {
      de[(size_t)i * ld + idx] = S.d[i];
      de[(size_t)(nmx + i) * ld + idx] = S.e[i];
      bws[(size_t)idx * nmx + i] = S.beta[i];
    }
    __syncthreads();
  }
}

template <typename T, int NCH, int CW, int RPL>
struct alignas(16) ApplyLargeLds {
  alignas(16) T red[2][NCH][RPL][64];            // per-wave partial dot products (phase A)
  alignas(16) T topv[NCH][RPL][64];              // wave q's top column after its last sweep
  alignas(16) T carry[NCH][RPL][64];             // pending chain value handed to wave q
  alignas(16) T us[NCH][CW];                     // reflector entries of wave q's columns
  alignas(16) typename vec2<T>::type rs[NCH][CW];  // rotations of wave q's current sweep
  alignas(16) int pos[NCH][CW];
  alignas(16) T beta[GL_NMAX];
};

// occupancy floor of the apply kernel (waves per SIMD): held to 128 VGPRs
template <typename T, int NCH> struct ApWaves { static constexpr int value = 4; };

// LIST: the uneven-batch form (u), one workgroup per (slot, row block of
// order nmx); otherwise (matrix, row block) of a uniform batch of order n.
// SKIPQ: start from the identity instead of Q (phase A skipped): the result is
// the tridiagonal's eigenvector matrix Z, for a separate back-transformation
// (eig_wy.hpp).
// RPL rows per lane: a workgroup takes a block of 64 RPL rows, lane l of every
// wave the rows 64 rr + l (rr < RPL) of the block.  The rows share each
// rotation's LDS read and run as RPL independent chains, so a sweep's chain
// through the wave's CW columns is RPL times as wide and no longer.
template <typename T, int NCH, int CW, bool LIST = false, bool SKIPQ = false, int RPL = 1>
__global__ __launch_bounds__(64 * NCH) __attribute__((amdgpu_waves_per_eu(ApWaves<T, NCH>::value))) void eig_apply_large_kernel(
    int L, int nm, int n, T *__restrict__ a, const T *__restrict__ W, const T *__restrict__ bws,
    const typename vec2<T>::type *__restrict__ rotlog, int rot_cap, const int *__restrict__ swlog, int sw_cap,
    const int *__restrict__ nsw_in, const int *__restrict__ status, const int *__restrict__ slot,
    UvSplitArgs u = UvSplitArgs{})
{
  typedef typename vec2<T>::type T2;
  __shared__ ApplyLargeLds<T, NCH, CW, RPL> S;
  constexpr int RB = 64 * RPL;                       // rows per workgroup
  const int nrb = ((LIST ? u.nmx : n) + RB - 1) / RB;
  const long long mat = blockIdx.x / nrb;            // the slot (LIST) or the matrix
  const int rb = blockIdx.x - (int)(mat * nrb);
  size_t wstride = (size_t)n * n, bstride = n, sstride = n;
  T *__restrict__ A;
  if constexpr (LIST) {
    int mi;
    if (!uv_split_entry(u, mat, mi)) return;
    n = __builtin_amdgcn_readfirstlane(u.nv[mi]);
    nm = n;
    if (rb * RB >= n) return;
    wstride = (size_t)u.nmx * u.nmx;
    bstride = sstride = u.nmx;
    A = a + u.aoff[mi];
  } else {
    if (mat >= L) return;
    A = a + (size_t)mat * (size_t)nm * (size_t)n;
  }
  if (status[mat] != 0) return;                      // overflow: solved by the fallback
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int c0 = q * CW;                             // first column of this wave
  const T *__restrict__ Wm = W + (size_t)mat * wstride;
  const T *__restrict__ beta = bws + (size_t)mat * bstride;
  int r[RPL];                                        // rows of Q handled by this lane
#pragma unroll
  for (int rr = 0; rr < RPL; ++rr) r[rr] = rb * RB + rr * 64 + lane;

  T x[RPL][CW];
#pragma unroll
  for (int rr = 0; rr < RPL; ++rr)
#pragma unroll
    for (int j = 0; j < CW; ++j) x[rr][j] = (r[rr] == c0 + j) ? T(1) : T(0);

  // ---- A. rows of Q = H_{n-1} ... H_1: x <- x H_c, c = n-1 down to 1 --------
  // (the right-multiplication of wave_front, split over the column chunks)
  for (int i = threadIdx.x; i < n; i += 64 * NCH) S.beta[i] = beta[i];
  // reflector entries are prefetched two reflectors ahead
  T unext = (lane < CW && c0 + lane < n - 1) ? Wm[c0 + lane + (size_t)(n - 1) * n] : T(0);
  T unext2 = (lane < CW && n >= 3 && c0 + lane < n - 2) ? Wm[c0 + lane + (size_t)(n - 2) * n] : T(0);
  __syncthreads();
  for (int c = SKIPQ ? 0 : n - 1; c >= 1; --c) {
    const T ucur = (c0 + lane < c) ? unext : T(0);
    unext = unext2;
    if (c > 2 && lane < CW) unext2 = (c0 + lane < c - 2) ? Wm[c0 + lane + (size_t)(c - 2) * n] : T(0);
    T part[RPL];
#pragma unroll
    for (int rr = 0; rr < RPL; ++rr) part[rr] = T(0);
    if (c0 < c) {
      if (lane < CW) S.us[q][lane] = ucur;
      wsync();
#pragma unroll
      for (int j = 0; j < CW; ++j) {
        const T uj = S.us[q][j];
#pragma unroll
        for (int rr = 0; rr < RPL; ++rr) part[rr] = fma_(x[rr][j], uj, part[rr]);
      }
    }
#pragma unroll
    for (int rr = 0; rr < RPL; ++rr) S.red[c & 1][q][rr][lane] = part[rr];
    __syncthreads();
    T tsum[RPL];
#pragma unroll
    for (int rr = 0; rr < RPL; ++rr) tsum[rr] = T(0);
    // a few waves' partials in flight at a time (all NCH RPL of them at once
    // would take the registers the rows need)
#pragma unroll 4
    for (int w2 = 0; w2 < NCH; ++w2)
#pragma unroll
      for (int rr = 0; rr < RPL; ++rr) tsum[rr] += S.red[c & 1][w2][rr][lane];
#pragma unroll
    for (int rr = 0; rr < RPL; ++rr) tsum[rr] *= S.beta[c];
    if (c0 < c) {
#pragma unroll
      for (int j = 0; j < CW; ++j) {
        const T uj = S.us[q][j];
#pragma unroll
        for (int rr = 0; rr < RPL; ++rr) x[rr][j] = fma_(tsum[rr], uj, x[rr][j]);
      }
    }
    wsync();
  }

  // ---- B. the logged sweeps, wave-pipelined over the column chunks ----------
  const T2 *__restrict__ lg = rotlog + (size_t)mat * rot_cap;
  const int *__restrict__ sw = swlog + (size_t)mat * sw_cap;
  const int nsw = nsw_in[mat];
  // this wave's pairs: plo .. phi (pair i finalises column i+1)
  const int phi = c0 + CW - 2;
  const int plo = (c0 - 1 > 0) ? c0 - 1 : 0;
  // Software pipeline: the rotations of the wave's next sweep are loaded
  // during the current one (lane j: pair c0-1+j, identity outside the sweep's
  // window), the sweep code one sweep further ahead.
  auto fetch = [&](int cd, int bs) -> T2 {
    const int m1 = sw_m(cd), ilo1 = sw_ilo(cd);
    const int i = c0 - 1 + lane;
    const bool in = lane < CW && i >= plo && i <= phi && i >= ilo1 && i <= m1 - 1;
    return in ? lg[bs + (m1 - 1 - i)] : T2{T(1), T(0)};
  };
  int base = 0;                                      // log offset of the wave's next sweep
  int code = (nsw > 0) ? sw[0] : 0;
  int code1 = (nsw > 1) ? sw[1] : 0;
  T2 pre = (nsw > 0) ? fetch(code, 0) : T2{T(1), T(0)};
#pragma unroll
  for (int rr = 0; rr < RPL; ++rr) S.topv[q][rr][lane] = x[rr][CW - 1];
  __syncthreads();
  const int nsteps = (nsw > 0) ? 2 * (nsw - 1) + NCH : 0;
  int s = 0;                                         // next sweep of this wave
  for (int step = 0; step < nsteps; ++step) {
    const int rel = step - (NCH - 1 - q);
    if (rel >= 0 && (rel & 1) == 0 && s < nsw) {
      const int m = sw_m(code), ilo = sw_ilo(code);
      const int cnt = m - ilo;
      const int top = (m - 1 < phi) ? m - 1 : phi;
      const int bot = (ilo > plo) ? ilo : plo;
      const bool enters = (m - 1 > phi) && (ilo <= phi + 1);
      if (lane < CW) S.rs[q][lane] = pre;
      const int code2 = (s + 2 < nsw) ? sw[s + 2] : 0;
      pre = (s + 1 < nsw) ? fetch(code1, base + cnt) : T2{T(1), T(0)};
      wsync();
      // 4-pair blocks from the one holding the window's top pair down to
      // the one holding its bottom pair; the rotations of a block are read
      // together from LDS as the block starts, and the identity rotations
      // outside the window make every pair of a block safe to apply.
      if (enters || top >= bot) {
        constexpr int PB = 4;
        const int ptop = enters ? CW - 1 : top - (c0 - 1);   // highest pair index p to apply
        const int pbot = bot - (c0 - 1);                      // lowest pair index p in the window
        const int btop = ptop / PB, bbot = pbot / PB;
        T h1[RPL], xlow[RPL];
#pragma unroll
        for (int rr = 0; rr < RPL; ++rr) {
          h1[rr] = T(0);
          if (enters) {
            h1[rr] = S.carry[q][rr][lane];
          } else {
#pragma unroll
            for (int bb = 0; bb < CW / PB; ++bb)                // x[btop PB + PB - 1]
              if (bb == btop) h1[rr] = x[rr][bb * PB + PB - 1];
          }
          xlow[rr] = (q > 0) ? S.topv[q - 1][rr][lane] : T(0);
        }
        T2 g[PB];
#pragma unroll
        for (int b = CW / PB - 1; b >= 0; --b) {
          if (b <= btop && b >= bbot) {
#pragma unroll
            for (int jj = 0; jj < PB; ++jj) g[jj] = S.rs[q][b * PB + jj];
#pragma unroll
            for (int jj = PB - 1; jj >= 0; --jj) {
              const int p = b * PB + jj;
#pragma unroll
              for (int rr = 0; rr < RPL; ++rr) {
                if (p >= 1) {
                  const T h0 = x[rr][p >= 1 ? p - 1 : 0];
                  x[rr][p] = g[jj].y * h0 + g[jj].x * h1[rr];
                  h1[rr] = g[jj].x * h0 - g[jj].y * h1[rr];
                } else if (q > 0) {
                  x[rr][0] = g[jj].y * xlow[rr] + g[jj].x * h1[rr];
                  h1[rr] = g[jj].x * xlow[rr] - g[jj].y * h1[rr];
                }
              }
            }
          }
        }
        // where the chain leaves this wave: column c0-1 (wave q-1) takes the
        // pending value; otherwise, if the window's bottom pair was the lowest
        // pair applied, its pending value is the final column ilo
        const int plow = (bbot * PB > (q > 0 ? 0 : 1)) ? bbot * PB : (q > 0 ? 0 : 1);
        if (q > 0 && pbot == 0) {
#pragma unroll
          for (int rr = 0; rr < RPL; ++rr) S.carry[q - 1][rr][lane] = h1[rr];
        } else if (plow == pbot && bot == ilo) {
#pragma unroll
          for (int j = 0; j < CW; ++j)
            if (j == pbot - 1) {
#pragma unroll
              for (int rr = 0; rr < RPL; ++rr) x[rr][j] = h1[rr];
            }
        }
      }
#pragma unroll
      for (int rr = 0; rr < RPL; ++rr) S.topv[q][rr][lane] = x[rr][CW - 1];
      base += cnt;
      ++s;
      code = code1;
      code1 = code2;
    }
    __syncthreads();
  }

  // ---- C. permuted store: column j goes to slot[j] --------------------------
  if (lane < CW) S.pos[q][lane] = (c0 + lane < n) ? slot[(size_t)mat * sstride + c0 + lane] : 0;
  wsync();
#pragma unroll
  for (int rr = 0; rr < RPL; ++rr) {
    if (r[rr] < n) {
#pragma unroll
      for (int j = 0; j < CW; ++j)
        if (c0 + j < n && S.pos[q][j] >= 0) A[r[rr] + (size_t)S.pos[q][j] * nm] = x[rr][j];
    }
  }
}

} // namespace eigmi
