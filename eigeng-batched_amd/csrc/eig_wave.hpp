// eig_wave.hpp - one matrix per group of NB lanes (n <= NB <= 64), whole
// pipeline in one kernel: Householder tridiagonalisation -> explicit
// Q = H_n ... H_2 -> QL with rotations accumulated into Q -> ascending sort ->
// permuted store of the eigenvectors.
//
// Layout (per group): lane k owns row k of the matrix in registers (A[NB], then
// Z[NB]; column index = register index, so every loop over columns is fully
// unrolled and column ranges are skipped in uniform blocks of 8).  The
// Householder vectors go to LDS packed by column (column c padded to a multiple
// of 8 with zeros) and are read back as 16-byte broadcasts; d, e, the rotation
// pairs (c, s) of one QL sweep and the sort permutation live in LDS too.
// Applying H_c from the right (Z <- Z H_c) and a Givens rotation of columns
// (i, i+1) are both lane-local in this layout, so no cross-lane traffic is
// needed outside the reductions of the tridiagonalisation.
//
// Reference path replaced: eigen_GPU_batch_tiled_ (eigen_GPU_batch.cu:56-135)
// with tred1_tiled_ (tred1_tiled.hpp:3-222), imtql2_tiled_/tql2_tiled_
// (imtql2_tiled.hpp:3-177, tql2_tiled.hpp:3-192) and hhtr2sy_tiled_
// (hhtr2sy_tiled.hpp:3-226).  The reference back-transforms after QL
// (Z <- H_2 ... H_n Z_QL); here Q = H_n ... H_2 is formed first and the QL
// rotations are accumulated into it, which gives the same product Q * Z_QL.
#pragma once
#include "common.hpp"

namespace eigmi {

enum { SOLVER_IMTQL2 = 0, SOLVER_TQL2 = 1, SOLVER_JACOBI = 2 };
constexpr int QL_MAX_SWEEP = 100; // tql2_tiled.hpp:7

// Offset of packed Householder column c (1 <= c) when column c' holds
// roundup8(c') entries: sum_{c'=1}^{c-1} 8*ceil(c'/8).
__host__ __device__ constexpr int uoff(int c)
{
  return 8 * ((c - 1) / 8 + 1) * (4 * ((c - 1) / 8) + (c - 1) % 8);
}

template <typename T, int NB>
struct alignas(16) WaveLds {
  alignas(16) T U[uoff(NB)];                 // packed Householder vectors
  alignas(16) T rowbuf[NB];                  // row c of A (= column c)
  alignas(16) T vbuf[NB];                    // v of the current step
  alignas(16) typename vec2<T>::type cs[NB]; // (c, s) of pair (i, i+1) in one QL sweep
  alignas(16) int pos[NB];                   // output slot of register column j
};

template <typename T, int NB>
__device__ __forceinline__ void apply_sweep_rotations(T (&Zr)[NB], const WaveLds<T, NB> &S, int lo, int hi)
{
  // Pairs (i, i+1) for i in [lo, hi], identity outside the sweep's own range,
  // applied from the top: h1 carries the pending value of column i+1.
  const int blo = lo >> 3, bhi = hi >> 3;
  T h1 = T(0);
#pragma unroll
  for (int b = NB / 8 - 1; b >= 0; --b) {
    if (b >= blo && b <= bhi) {
      const int itop = (8 * b + 7 < NB - 1) ? 8 * b + 7 : NB - 2;
      if (b == bhi) h1 = Zr[itop + 1];
#pragma unroll
      for (int i = itop; i >= 8 * b; --i) {
        const typename vec2<T>::type q = S.cs[i];
        const T h0 = Zr[i];
        Zr[i + 1] = q.y * h0 + q.x * h1;
        h1 = q.x * h0 - q.y * h1;
      }
      if (b == blo) Zr[8 * b] = h1;
    }
  }
}

// Set element i (group-uniform) of a lane-distributed vector.
template <typename T>
__device__ __forceinline__ T gset(T v, int k, int i, T x) { return (k == i) ? x : v; }

template <typename T, int NB>
__global__ __launch_bounds__(64) void eig_wave_kernel(int L, int nm, int n, T *__restrict__ a,
                                                      T *__restrict__ w, int *__restrict__ info,
                                                      int solver)
{
  constexpr int G = 64 / NB;
  constexpr int NBLK = NB / 8;
  typedef typename vec2<T>::type T2;
  __shared__ WaveLds<T, NB> smem[G];

  const int lane = threadIdx.x & 63;
  const int g = lane / NB, k = lane % NB;
  const long long mat = (long long)blockIdx.x * G + g;
  if (mat >= L) return;  // whole group leaves together
  WaveLds<T, NB> &S = smem[g];
  T *__restrict__ A = a + (size_t)mat * (size_t)nm * (size_t)n;

  // R holds row k of A during the reduction and row k of Z afterwards.
  T R[NB];
  // ---- load: row k of the symmetric matrix from the upper triangle ----------
  // Columns j > k come from row k of the upper triangle (coalesced across
  // lanes); columns j <= k come from column k (a(j,k), j <= k, is the upper
  // triangle too), read by each lane as one contiguous run with 16-byte loads.
  {
    const int kc = (k < n) ? k : n - 1;
    constexpr int V = 16 / sizeof(T);      // elements per 16-byte load
    typedef T vload_t __attribute__((ext_vector_type(V), aligned(sizeof(T))));
    const T *colk = A + (size_t)kc * nm;
#pragma unroll
    for (int q = 0; q < NB / V; ++q) {
      T v[V];
      if (V * q + V <= n) {
        const vload_t x = *(const vload_t *)(colk + V * q);
#pragma unroll
        for (int t = 0; t < V; ++t) v[t] = x[t];
      } else {
#pragma unroll
        for (int t = 0; t < V; ++t) v[t] = (V * q + t < n) ? colk[V * q + t] : T(0);
      }
#pragma unroll
      for (int t = 0; t < V; ++t) R[V * q + t] = v[t];
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int jc = (j < n) ? j : n - 1;
      const T v = A[(size_t)kc + (size_t)jc * nm];
      R[j] = (j < n && k < n) ? ((j > k) ? v : R[j]) : T(0);
    }
  }

  // d, e and beta are lane-distributed: element c lives in lane c of the group.
  T dreg = T(0), ereg = T(0), breg = T(0);

  // ---- Householder tridiagonalisation (tred1_tiled.hpp:42-212) --------------
  for (int c = n - 1; c >= 1; --c) {
    const int nbr = (c >> 3) + 1;           // 8-blocks covering columns 0..c
    if (k == c) {
#pragma unroll
      for (int b = 0; b < NBLK; ++b)
        if (b < nbr) st8(&S.rowbuf[8 * b], &R[8 * b]);
    }
    wsync();
    T uk = (k < c) ? S.rowbuf[k] : T(0);
    const T dc = S.rowbuf[c];
    const T ulast = S.rowbuf[c - 1];
    const T sig = group_sum<NB>(uk * uk);
    const T anorm = -sign_(sqrt_(sig), ulast);
    const T ul = ulast - anorm;
    const T beta = T(1) / nz(ul * anorm);
    if (k == c - 1) uk = ul;
    const int off = uoff(c);
    const int nbc = (c + 7) >> 3;           // 8-blocks covering columns 0..c-1
    if (k < 8 * nbc) S.U[off + k] = uk;
    dreg = gset(dreg, k, c, dc);
    ereg = gset(ereg, k, c - 1, anorm);
    breg = gset(breg, k, c, beta);
    wsync();
    // p = beta * A(0:c-1, 0:c-1) u  (row k of the product)
    T p = T(0);
#pragma unroll
    for (int b = 0; b < NBLK; ++b) {
      if (b < nbc) {
        T u8[8];
        ld8(&S.U[off + 8 * b], u8);
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) p = fma_(R[8 * b + jj], u8[jj], p);
      }
    }
    p = (k < c) ? p * beta : T(0);
    const T alpha = group_sum<NB>(p * uk) * (beta * T(0.5));
    const T vk = fma_(alpha, uk, p);
    S.vbuf[k] = vk;
    wsync();
    // A(0:c-1,0:c-1) += u v^T + v u^T
#pragma unroll
    for (int b = 0; b < NBLK; ++b) {
      if (b < nbc) {
        T u8[8], v8[8];
        ld8(&S.U[off + 8 * b], u8);
        ld8(&S.vbuf[8 * b], v8);
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) R[8 * b + jj] = fma_(uk, v8[jj], fma_(vk, u8[jj], R[8 * b + jj]));
      }
    }
    wsync();
  }
  dreg = gset(dreg, k, 0, R[0]);

  // ---- Q = H_{n-1} ... H_1 (0-based creation order), right-multiplied -------
#pragma unroll
  for (int j = 0; j < NB; ++j) R[j] = (j == k) ? T(1) : T(0);
  for (int c = n - 1; c >= 1; --c) {
    const int off = uoff(c);
    const int nbc = (c + 7) >> 3;
    const T beta = gread<NB>(breg, c);
    T t = T(0);
#pragma unroll
    for (int b = 0; b < NBLK; ++b) {
      if (b < nbc) {
        T u8[8];
        ld8(&S.U[off + 8 * b], u8);
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) t = fma_(R[8 * b + jj], u8[jj], t);
      }
    }
    t *= beta;
#pragma unroll
    for (int b = 0; b < NBLK; ++b) {
      if (b < nbc) {
        T u8[8];
        ld8(&S.U[off + 8 * b], u8);
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) R[8 * b + jj] = fma_(t, u8[jj], R[8 * b + jj]);
      }
    }
  }

  // ---- QL with eigenvectors accumulated into Z = R --------------------------
  // (c, s) of the current sweep: lane i holds the pair (i, i+1); identity
  // outside the sweep.  Dumped to LDS once per sweep for the rotation chain.
  const T tol = eps_of<T>() * T(sizeof(T) == 8 ? 512 : 16);
  T csc = T(1), css = T(0);
  int err = 0;
  if (solver == SOLVER_TQL2) {
    // explicit shift (tql2_tiled.hpp:40-146)
    T shift = T(0), tst1 = T(0);
    for (int l = 0; l < n && !err; ++l) {
      const T h = (abs_(gread<NB>(dreg, l)) + abs_(gread<NB>(ereg, l))) * tol;
      tst1 = (h > tst1) ? h : tst1;
      const bool cand = (k >= l && k <= n - 1) && (abs_(ereg) <= tst1);
      const uint64_t msk = group_ballot<NB>(cand);
      const int m = msk ? __builtin_ctzll(msk) : n - 1;
      if (m != l) {
        int itr;
        for (itr = 0; itr < QL_MAX_SWEEP; ++itr) {
          const T dl_old = gread<NB>(dreg, l), el = gread<NB>(ereg, l);
          const T p0 = (gread<NB>(dreg, l + 1) - dl_old) / (el + el);
          const T r0 = hypot1_c(p0);
          const T psr = p0 + sign_(r0, p0);
          const T dl = el / psr, dl1 = el * psr;
          const T delta = dl_old - dl;
          dreg = (k == l) ? dl : (k == l + 1) ? dl1 : (k >= l + 2 && k < n) ? dreg - delta : dreg;
          shift += delta;
          T c = T(1), c2 = T(1), c3 = T(1), s = T(0), s2 = T(0);
          T pp = gread<NB>(dreg, m);
          const T el1 = gread<NB>(ereg, l + 1);
          T ei = gread<NB>(ereg, m - 1), di = gread<NB>(dreg, m - 1);
          for (int i = m - 1; i >= l; --i) {
            const int ip = (i > 0) ? i - 1 : 0;
            const T ein = gread<NB>(ereg, ip), din = gread<NB>(dreg, ip);
            c3 = c2; c2 = c; s2 = s;
            const T gg = c * ei, hh = c * pp;
            T r, cn, sn;
            if (!givens(ei, pp, r, cn, sn)) { cn = T(1); sn = T(0); }
            const T ei1 = s * r;
            s = sn;
            c = cn;
            pp = c * di - s * gg;
            const T di1 = hh + s * (c * gg + s * di);
            ereg = gset(ereg, k, i + 1, ei1);
            dreg = gset(dreg, k, i + 1, di1);
            csc = gset(csc, k, i, c);
            css = gset(css, k, i, s);
            ei = ein;
            di = din;
          }
          const T rr = (-s * s2 * c3 * el1 * el) / dl1;
          const T el_new = s * rr;
          ereg = gset(ereg, k, l, el_new);
          dreg = gset(dreg, k, l, c * rr);
          S.cs[k] = T2{csc, css};
          wsync();
          apply_sweep_rotations<T, NB>(R, S, l, m - 1);
          wsync();
          csc = T(1);
          css = T(0);
          if (abs_(el_new) <= tst1) break;
        }
        if (itr >= QL_MAX_SWEEP) err = l + 1;
      }
      if (!err) dreg = (k == l) ? dreg + shift : dreg;
    }
  } else {
    // implicit shift (imtql2_tiled.hpp:235-321)
    for (int l = 0; l < n && !err; ++l) {
      for (int itr = 0;; ++itr) {
        // itr sweeps done; the reference stops after max_sweep+1 sweeps without
        // a final test (imtql2_tiled.hpp:242,319)
        if (itr > QL_MAX_SWEEP) { err = l + 1; break; }
        const T dk1 = lane_bcast(dreg, lane + 1);
        const bool cand = (k >= l && k <= n - 2) && abs_(ereg) <= (abs_(dreg) + abs_(dk1)) * tol;
        const uint64_t msk = group_ballot<NB>(cand);
        const int m = msk ? __builtin_ctzll(msk) : n - 1;
        if (m == l) break;
        const T dl = gread<NB>(dreg, l), el = gread<NB>(ereg, l);
        const T f0 = (gread<NB>(dreg, l + 1) - dl) / (el + el);
        const T g0 = hypot1_c(f0);
        T di1 = gread<NB>(dreg, m);
        T r = (di1 - dl) + el / (f0 + sign_(g0, f0));
        T s = T(1), c = T(1), delta = T(0);
        T ei = gread<NB>(ereg, m - 1), di = gread<NB>(dreg, m - 1);
        int i;
        for (i = m - 1; i >= l; --i) {
          const int ip = (i > 0) ? i - 1 : 0;
          const T ein = gread<NB>(ereg, ip), din = gread<NB>(dreg, ip);
          const T f = s * ei, gg = r;
          const T bb = c * ei;                   // old c
          if (!givens(f, gg, r, c, s)) break;    // r = 0: underflow, restart the sweep
          const T dix = di1 - delta;
          di1 = di;
          const T q = (di1 - dix) * s + T(2) * c * bb;
          const T dx1 = fma_(s, q, dix);
          delta = dx1 - dix;
          ereg = gset(ereg, k, i + 1, r);
          dreg = gset(dreg, k, i + 1, dx1);
          csc = gset(csc, k, i, c);
          css = gset(css, k, i, s);
          r = c * q - bb;
          ei = ein;
          di = din;
        }
        const int j = (i + 1 > l) ? i + 1 : l;
        dreg = (k == j) ? dreg - delta : dreg;
        ereg = gset(ereg, k, j, r);
        ereg = gset(ereg, k, m, T(0));
        S.cs[k] = T2{csc, css};
        wsync();
        apply_sweep_rotations<T, NB>(R, S, l, m - 1);
        wsync();
        csc = T(1);
        css = T(0);
      }
    }
  }

  if (err) {
    if (k < n) w[(size_t)mat * n + k] = dreg;
    if (info && k == 0) info[mat] = err;
    return;
  }
  if (info && k == 0) info[mat] = 0;

  // ---- ascending sort, the reference's selection rule (tql2_tiled.hpp:149-176):
  // step l swaps slot l with the first minimum of slots l..n-1.
  T dk = dreg;
  int pk = k;
  const int base = lane & ~(NB - 1);
  for (int l = 0; l < n - 1; ++l) {
    const bool in = (k >= l && k < n);
    T v = dk;
    int ix = in ? k : NB;
    // argmin (value, index), ties to the smaller index; out-of-range lanes lose
#pragma unroll
    for (int sh = 1; sh < NB; sh <<= 1) {
      const T v2 = lane_bcast(v, lane ^ sh);
      const int i2 = lane_bcast(ix, lane ^ sh);
      const bool take = (i2 < NB) && (ix == NB || v2 < v || (v2 == v && i2 < ix));
      v = take ? v2 : v;
      ix = take ? i2 : ix;
    }
    if (ix != l) {
      const T d_l = lane_bcast(dk, base + l);
      const int p_l = lane_bcast(pk, base + l);
      const int p_i = lane_bcast(pk, base + ix);
      if (k == l) { dk = v; pk = p_i; }
      if (k == ix) { dk = d_l; pk = p_l; }
    }
  }
  if (k < n) {
    w[(size_t)mat * n + k] = dk;
    S.pos[pk] = k;
  }
  wsync();
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < n) {
      const int r = S.pos[j];
      if (k < n) A[k + (size_t)r * nm] = R[j];
    }
  }
}

} // namespace eigmi
