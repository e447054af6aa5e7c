// eig_wave.hpp - register-resident solver for n <= NB <= 128: one matrix per
// group of NB lanes (NB = 8..64: a group inside one wave64; NB = 128: two
// waves), whole pipeline in one kernel: Householder tridiagonalisation ->
// explicit Q = H_n ... H_2 -> QL with the rotations accumulated into Q ->
// ascending sort -> permuted store of the eigenvectors.
//
// Layout: lane k of the group owns row k of the matrix in registers (R[NB],
// column index = register index, so every loop over columns is fully unrolled
// and column ranges are skipped in uniform blocks of 8).  R holds A during the
// reduction and Z afterwards.  The Householder vectors go to LDS packed by
// column (column c padded to a multiple of 8 with zeros) and are read back as
// 16-byte broadcasts.  Applying H_c from the right (Z <- Z H_c) and a Givens
// rotation of columns (i, i+1) are both lane-local in this layout.
//
// The tridiagonal (d, e), the betas and the (c, s) of the current QL sweep are
// lane-distributed vectors: element i lives in lane i % 64 of every wave of the
// group (slot i / 64), replicated in each wave, so the QL phase runs with no
// cross-wave synchronisation at all; for a single wave, element i is read with
// v_readlane (wave-uniform index in an SGPR), otherwise with ds_bpermute.
//
// Reference path replaced: eigen_GPU_batch_tiled_ (eigen_GPU_batch.cu:56-135)
// with tred1_tiled_ (tred1_tiled.hpp:3-222), imtql2_tiled_/tql2_tiled_
// (imtql2_tiled.hpp:3-177, tql2_tiled.hpp:3-192) and hhtr2sy_tiled_
// (hhtr2sy_tiled.hpp:3-226); for 64 < n <= 128 the reference's wave kernel
// eigen_GPU_batch_ (eigen_GPU_batch.cu:138-245).  The reference back-transforms
// after QL (Z <- H_2 ... H_n Z_QL); here Q = H_n ... H_2 is formed first and the
// QL rotations are accumulated into it, which gives the same product Q * Z_QL.
#pragma once
#include "common.hpp"

namespace eigmi {

enum { SOLVER_IMTQL2 = 0, SOLVER_TQL2 = 1, SOLVER_JACOBI = 2 };
constexpr int QL_MAX_SWEEP = 100; // tql2_tiled.hpp:7
constexpr int ST_OVERFLOW = -2;
// Output options packed with the solver id into the kernels' `solver`
// argument: bits 0-7 the solver, bit 8 keep the QL / Jacobi order of the
// eigenpairs (the reference's DO_SORT 0, eigen_GPU_batch.cu:5,10), bits 12-21
// the number of trailing eigenvector columns not written (n - m).  Launches
// that pass a bare solver id get the default: sorted, all n columns.
__host__ __device__ constexpr int kopt_make(int solver, bool unsorted, int mdrop)
{
  return (solver & 0xff) | (unsorted ? 0x100 : 0) | ((mdrop & 0x3ff) << 12);
}
__host__ __device__ constexpr int kopt_solver(int o) { return o & 0xff; }
__host__ __device__ constexpr bool kopt_unsorted(int o) { return ((o >> 8) & 1) != 0; }
__host__ __device__ constexpr int kopt_mdrop(int o) { return (o >> 12) & 0x3ff; }   // split path: rotation log full, solve with the fused kernel

// Offset of packed Householder column c (1 <= c) when column c' holds
// roundup8(c') entries: sum_{c'=1}^{c-1} 8*ceil(c'/8).
__host__ __device__ constexpr int uoff(int c)
{
  return 8 * ((c - 1) / 8 + 1) * (4 * ((c - 1) / 8) + (c - 1) % 8);
}

template <int NB> struct GroupShape {
  static constexpr int NL = NB < 64 ? NB : 64; // lanes of the group inside one wave
  static constexpr int NW = NB / NL;           // waves per group
  static constexpr int G = 64 / NL;            // groups per wave
};

template <typename T, int NB>
struct alignas(16) WaveLds {
  static constexpr int NW = GroupShape<NB>::NW;
  alignas(16) T U[uoff(NB)];                     // packed Householder vectors
  alignas(16) T rowbuf[NB];                      // row c of A (= column c)
  alignas(16) T vbuf[NB];                        // v of the current step
  alignas(16) T red[2][2];                       // cross-wave partial sums
  alignas(16) typename vec2<T>::type cs[NW][NB]; // per wave: (c, s) of pair (i, i+1)
  alignas(16) int pos[NW][NB];                   // per wave: output slot of column j
};

template <int NW>
__device__ __forceinline__ void gsync()
{
  if (NW == 1) wsync();
  else __syncthreads();
}

// Sum over the NB lanes of the group; `slot` selects one of two scratch pairs
// (two consecutive reductions must use different slots or be separated by a
// group barrier).
template <int NB, typename T>
__device__ __forceinline__ T gsum(T x, T (&red)[2][2], int slot, int wv, int kl)
{
  constexpr int NL = GroupShape<NB>::NL, NW = GroupShape<NB>::NW;
  x = group_sum<NL>(x);
  if (NW == 1) return x;
  if (kl == 0) red[slot][wv] = x;
  __syncthreads();
  return red[slot][0] + red[slot][1];
}

// Lane-distributed vectors: element i in slot i / NL of lane i % NL.
template <int NL, int NW, typename T>
__device__ __forceinline__ T vget(const T (&v)[NW], int i)
{
  if (NW == 1) return gread<NL>(v[0], i);
  return (i < NL) ? gread<NL>(v[0], i) : gread<NL>(v[NW - 1], i - NL);
}
template <int NL, int NW, typename T>
__device__ __forceinline__ void vset(T (&v)[NW], int kl, int i, T x)
{
#pragma unroll
  for (int h = 0; h < NW; ++h) v[h] = (kl + h * NL == i) ? x : v[h];
}

template <typename T, int NB>
__device__ __forceinline__ void apply_sweep_rotations(T (&Zr)[NB], const typename vec2<T>::type *cs,
                                                      int lo, int hi)
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
        const typename vec2<T>::type q = cs[i];
        const T h0 = Zr[i];
        Zr[i + 1] = q.y * h0 + q.x * h1;
        h1 = q.x * h0 - q.y * h1;
      }
      if (b == blo) Zr[8 * b] = h1;
    }
  }
}

// Two consecutive sweeps A then B, B one 8-block behind A: in round b, A's
// pairs of block b and then B's pairs of block b+1.  B's pair j needs columns
// j and j+1 as A leaves them, which A's pairs j-1 and j (blocks at or above
// b) have finished, so every element sees A's rotations before B's exactly
// as in two apply_sweep_rotations calls, and the two chains of a round sit in
// one basic block so that their latencies overlap.  A round runs both blocks
// whenever either is inside its sweep's window; a block outside its own
// window is a no-op: its rotations are the identity, and its chain restarts
// at the block's top and commits at its bottom.
template <typename T, int NB>
__device__ __forceinline__ void apply_two_sweeps(T (&Zr)[NB], const typename vec2<T>::type *csA, int loA, int hiA,
                                                 const typename vec2<T>::type *csB, int loB, int hiB)
{
  typedef typename vec2<T>::type T2;
  constexpr int NBLK = NB / 8;
  const int bloA = loA >> 3, bhiA = hiA >> 3, bloB = loB >> 3, bhiB = hiB >> 3;
  // rounds b with A's block b or B's block b+1 in its window
  const int rlo = (bloA < bloB - 1) ? bloA : bloB - 1, rhi = (bhiA > bhiB - 1) ? bhiA : bhiB - 1;
  T hA = T(0), hB = T(0);
  auto blk = [&](const T2 *cs, T &h1, int b, int blo, int bhi) {
    const int itop = (8 * b + 7 < NB - 1) ? 8 * b + 7 : NB - 2;
    h1 = (b >= bhi) ? Zr[itop + 1] : h1;
#pragma unroll
    for (int i = itop; i >= 8 * b; --i) {
      const T2 q = cs[i];
      const T h0 = Zr[i];
      Zr[i + 1] = q.y * h0 + q.x * h1;
      h1 = q.x * h0 - q.y * h1;
    }
    Zr[8 * b] = (b <= blo) ? h1 : Zr[8 * b];
  };
#pragma unroll
  for (int b = NBLK - 1; b >= -1; --b) {
    if (b >= rlo && b <= rhi) {
      // keeps each round's rotation loads in their round (hoisted ahead, the
      // loads of later rounds would hold far more registers than the row)
      asm volatile("" ::: "memory");
      if (b >= 0) blk(csA, hA, b >= 0 ? b : 0, bloA, bhiA);
      if (b + 1 <= NBLK - 1) blk(csB, hB, b + 1 <= NBLK - 1 ? b + 1 : NBLK - 1, bloB, bhiB);
    }
  }
}

// Front half shared by the fused and the split solvers: load row k of matrix A
// (upper triangle only), Householder tridiagonalisation, Q = H_{n-1}...H_1 in R.
// On return R holds row k of Q and dv / ev the lane-distributed diagonal and
// off-diagonal (replicated in every wave of the group).
template <typename T, int NB>
__device__ __forceinline__ void wave_front(const T *__restrict__ A, int nm, int n, int k, int kl, int wv,
                                           WaveLds<T, NB> &S, T (&R)[NB],
                                           T (&dv)[GroupShape<NB>::NW], T (&ev)[GroupShape<NB>::NW])
{
  constexpr int NL = GroupShape<NB>::NL, NW = GroupShape<NB>::NW;
  constexpr int NBLK = NB / 8;
  // R holds row k of A during the reduction and row k of Q afterwards.
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

  T bv[NW];
#pragma unroll
  for (int h = 0; h < NW; ++h) { dv[h] = T(0); ev[h] = T(0); bv[h] = T(0); }

  // ---- Householder tridiagonalisation (tred1_tiled.hpp:42-212) --------------
  for (int c = n - 1; c >= 1; --c) {
    const int nbr = (c >> 3) + 1;           // 8-blocks covering columns 0..c
    if (k == c) {
#pragma unroll
      for (int b = 0; b < NBLK; ++b)
        if (b < nbr) st8(&S.rowbuf[8 * b], &R[8 * b]);
    }
    gsync<NW>();
    T uk = (k < c) ? S.rowbuf[k] : T(0);
    const T dc = S.rowbuf[c];
    const T ulast = S.rowbuf[c - 1];
    const T sig = gsum<NB>(uk * uk, S.red, 0, wv, kl);
    const T anorm = -sign_(sqrt_(sig), ulast);
    const T ul = ulast - anorm;
    const T beta = T(1) / nz(ul * anorm);
    if (k == c - 1) uk = ul;
    const int off = uoff(c);
    const int nbc = (c + 7) >> 3;           // 8-blocks covering columns 0..c-1
    if (k < 8 * nbc) S.U[off + k] = uk;
    vset<NL, NW>(dv, kl, c, dc);
    vset<NL, NW>(ev, kl, c - 1, anorm);
    vset<NL, NW>(bv, kl, c, beta);
    gsync<NW>();
    // p = beta * A(0:c-1, 0:c-1) u  (row k of the product), in two partial
    // sums so that the FMA chain is half as long
    T p = T(0), p1 = T(0);
#pragma unroll
    for (int b = 0; b < NBLK; ++b) {
      if (b < nbc) {
        T u8[8];
        ld8(&S.U[off + 8 * b], u8);
#pragma unroll
        for (int jj = 0; jj < 8; jj += 2) {
          p = fma_(R[8 * b + jj], u8[jj], p);
          p1 = fma_(R[8 * b + jj + 1], u8[jj + 1], p1);
        }
      }
    }
    p += p1;
    p = (k < c) ? p * beta : T(0);
    const T alpha = gsum<NB>(p * uk, S.red, 1, wv, kl) * (beta * T(0.5));
    const T vk = fma_(alpha, uk, p);
    S.vbuf[k] = vk;
    gsync<NW>();
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
    if (NW == 1) wsync();
  }
  {
    // d(0) = A(0,0) lives in row 0 (lane 0 of wave 0); broadcast to every wave
    if (k == 0) S.rowbuf[0] = R[0];
    gsync<NW>();
    vset<NL, NW>(dv, kl, 0, S.rowbuf[0]);
  }

  // ---- Q = H_{n-1} ... H_1 (0-based creation order), right-multiplied -------
#pragma unroll
  for (int j = 0; j < NB; ++j) R[j] = (j == k) ? T(1) : T(0);
  for (int c = n - 1; c >= 1; --c) {
    const int off = uoff(c);
    const int nbc = (c + 7) >> 3;
    const T beta = vget<NL, NW>(bv, c);
    T t = T(0), t1 = T(0);
#pragma unroll
    for (int b = 0; b < NBLK; ++b) {
      if (b < nbc) {
        T u8[8];
        ld8(&S.U[off + 8 * b], u8);
#pragma unroll
        for (int jj = 0; jj < 8; jj += 2) {
          t = fma_(R[8 * b + jj], u8[jj], t);
          t1 = fma_(R[8 * b + jj + 1], u8[jj + 1], t1);
        }
      }
    }
    t = (t + t1) * beta;
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

}

// Whole fused solve of one matrix by its group: A (leading dimension nm, order
// n), eigenvalues to wout[0..n-1], status to *infop (may be null).
template <typename T, int NB>
__device__ __forceinline__ void wave_solve(T *__restrict__ A, int nm, int n, T *__restrict__ wout,
                                           int *__restrict__ infop, int solver, WaveLds<T, NB> &S,
                                           int lane, int wv, int kl, int k)
{
  constexpr int NL = GroupShape<NB>::NL, NW = GroupShape<NB>::NW;
  typedef typename vec2<T>::type T2;
  T R[NB], dv[NW], ev[NW];
  wave_front<T, NB>(A, nm, n, k, kl, wv, S, R, dv, ev);

  // ---- QL with eigenvectors accumulated into Z = R --------------------------
  // Everything below is per wave (replicated vectors, a private (c, s) array):
  // no cross-wave synchronisation.
  T2 *const cs = S.cs[wv];
  const T tol = eps_of<T>() * T(sizeof(T) == 8 ? 512 : 16);
  T csc[NW], css[NW];
#pragma unroll
  for (int h = 0; h < NW; ++h) { csc[h] = T(1); css[h] = T(0); }
  int err = 0;

  // first index m >= l with the predicate set (lane-distributed), else n - 1
  auto first_set = [&](const bool (&cand)[NW]) -> int {
    if (NW == 1) {
      const uint64_t msk = group_ballot<NL>(cand[0]);
      return msk ? __builtin_ctzll(msk) : n - 1;
    }
#pragma unroll
    for (int h = 0; h < NW; ++h) {
      const uint64_t msk = __ballot(cand[h]);
      if (msk) return h * 64 + __builtin_ctzll(msk);
    }
    return n - 1;
  };
  auto dump_cs = [&]() {
#pragma unroll
    for (int h = 0; h < NW; ++h) cs[kl + h * NL] = T2{csc[h], css[h]};
  };
  auto reset_cs = [&]() {
#pragma unroll
    for (int h = 0; h < NW; ++h) { csc[h] = T(1); css[h] = T(0); }
  };

  if (kopt_solver(solver) == SOLVER_TQL2) {
    // explicit shift (tql2_tiled.hpp:40-146)
    T shift = T(0), tst1 = T(0);
    for (int l = 0; l < n && !err; ++l) {
      const T h = (abs_(vget<NL, NW>(dv, l)) + abs_(vget<NL, NW>(ev, l))) * tol;
      tst1 = (h > tst1) ? h : tst1;
      bool cand[NW];
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        const int idx = kl + q * NL;
        cand[q] = (idx >= l && idx <= n - 1) && (abs_(ev[q]) <= tst1);
      }
      const int m = first_set(cand);
      if (m != l) {
        int itr;
        for (itr = 0; itr < QL_MAX_SWEEP; ++itr) {
          const T dl_old = vget<NL, NW>(dv, l), el = vget<NL, NW>(ev, l);
          const T p0 = (vget<NL, NW>(dv, l + 1) - dl_old) / (el + el);
          const T r0 = hypot1_c(p0);
          const T psr = p0 + sign_(r0, p0);
          const T dl = el / psr, dl1 = el * psr;
          const T delta = dl_old - dl;
#pragma unroll
          for (int q = 0; q < NW; ++q) {
            const int idx = kl + q * NL;
            dv[q] = (idx == l) ? dl : (idx == l + 1) ? dl1 : (idx >= l + 2 && idx < n) ? dv[q] - delta : dv[q];
          }
          shift += delta;
          T c = T(1), c2 = T(1), c3 = T(1), s = T(0), s2 = T(0);
          T pp = vget<NL, NW>(dv, m);
          const T el1 = vget<NL, NW>(ev, l + 1);
          T ei = vget<NL, NW>(ev, m - 1), di = vget<NL, NW>(dv, m - 1);
          for (int i = m - 1; i >= l; --i) {
            const int ip = (i > 0) ? i - 1 : 0;
            const T ein = vget<NL, NW>(ev, ip), din = vget<NL, NW>(dv, ip);
            c3 = c2; c2 = c; s2 = s;
            const T gg = c * ei, hh = c * pp;
            T r, cn, sn;
            if (!givens(ei, pp, r, cn, sn)) { cn = T(1); sn = T(0); }
            const T ei1 = s * r;
            s = sn;
            c = cn;
            pp = c * di - s * gg;
            const T di1 = hh + s * (c * gg + s * di);
            vset<NL, NW>(ev, kl, i + 1, ei1);
            vset<NL, NW>(dv, kl, i + 1, di1);
            vset<NL, NW>(csc, kl, i, c);
            vset<NL, NW>(css, kl, i, s);
            ei = ein;
            di = din;
          }
          const T rr = (-s * s2 * c3 * el1 * el) / dl1;
          const T el_new = s * rr;
          vset<NL, NW>(ev, kl, l, el_new);
          vset<NL, NW>(dv, kl, l, c * rr);
          dump_cs();
          wsync();
          apply_sweep_rotations<T, NB>(R, cs, l, m - 1);
          wsync();
          reset_cs();
          if (abs_(el_new) <= tst1) break;
        }
        if (itr >= QL_MAX_SWEEP) err = l + 1;
      }
      if (!err) {
#pragma unroll
        for (int q = 0; q < NW; ++q) dv[q] = (kl + q * NL == l) ? dv[q] + shift : dv[q];
      }
    }
  } else {
    // implicit shift (imtql2_tiled.hpp:43-134)
    for (int l = 0; l < n && !err; ++l) {
      for (int itr = 0;; ++itr) {
        // itr sweeps done; the reference stops after max_sweep+1 sweeps without
        // a final test (imtql2_tiled.hpp:49,126)
        if (itr > QL_MAX_SWEEP) { err = l + 1; break; }
        bool cand[NW];
#pragma unroll
        for (int q = 0; q < NW; ++q) {
          const int idx = kl + q * NL;
          // d(idx+1): next lane of the same slot, or lane 0 of the next slot
          T dn = lane_bcast(dv[q], lane + 1);
          if (NW > 1) {
            const T d0 = lane_bcast(dv[q + 1 < NW ? q + 1 : q], 0);
            dn = (kl == NL - 1) ? d0 : dn;
          }
          cand[q] = (idx >= l && idx <= n - 2) && abs_(ev[q]) <= (abs_(dv[q]) + abs_(dn)) * tol;
        }
        const int m = first_set(cand);
        if (m == l) break;
        const T dl = vget<NL, NW>(dv, l), el = vget<NL, NW>(ev, l);
        const T f0 = (vget<NL, NW>(dv, l + 1) - dl) / (el + el);
        const T g0 = hypot1_c(f0);
        T di1 = vget<NL, NW>(dv, m);
        T r = (di1 - dl) + el / (f0 + sign_(g0, f0));
        T s = T(1), c = T(1), delta = T(0);
        T ei = vget<NL, NW>(ev, m - 1), di = vget<NL, NW>(dv, m - 1);
        int i;
        for (i = m - 1; i >= l; --i) {
          const int ip = (i > 0) ? i - 1 : 0;
          const T ein = vget<NL, NW>(ev, ip), din = vget<NL, NW>(dv, ip);
          const T f = s * ei, gg = r;
          const T bb = c * ei;                   // old c
          if (!givens(f, gg, r, c, s)) break;    // r = 0: underflow, restart the sweep
          const T dix = di1 - delta;
          di1 = di;
          const T q = (di1 - dix) * s + T(2) * c * bb;
          const T dx1 = fma_(s, q, dix);
          delta = dx1 - dix;
          vset<NL, NW>(ev, kl, i + 1, r);
          vset<NL, NW>(dv, kl, i + 1, dx1);
          vset<NL, NW>(csc, kl, i, c);
          vset<NL, NW>(css, kl, i, s);
          r = c * q - bb;
          ei = ein;
          di = din;
        }
        const int j = (i + 1 > l) ? i + 1 : l;
#pragma unroll
        for (int q = 0; q < NW; ++q) dv[q] = (kl + q * NL == j) ? dv[q] - delta : dv[q];
        vset<NL, NW>(ev, kl, j, r);
        vset<NL, NW>(ev, kl, m, T(0));
        dump_cs();
        wsync();
        apply_sweep_rotations<T, NB>(R, cs, l, m - 1);
        wsync();
        reset_cs();
      }
    }
  }

  if (err) {
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      const int idx = kl + q * NL;
      if (q == wv && idx < n) wout[idx] = dv[q];
    }
    if (infop && k == 0) *infop = err;
    return;
  }
  if (infop && k == 0) *infop = 0;

  // ---- ascending sort, the reference's selection rule (tql2_tiled.hpp:149-176):
  // step l swaps slot l with the first minimum of slots l..n-1.
  int pv[NW];
#pragma unroll
  for (int q = 0; q < NW; ++q) pv[q] = kl + q * NL;
  const int nsort = kopt_unsorted(solver) ? 0 : n - 1;
  for (int l = 0; l < nsort; ++l) {
    // this lane's best candidate over its slots (ties to the smaller index)
    T v = T(0);
    int ix = NB;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      const int idx = kl + q * NL;
      const bool in = (idx >= l && idx < n);
      const bool take = in && (ix == NB || dv[q] < v);
      v = take ? dv[q] : v;
      ix = take ? idx : ix;
    }
    // argmin (value, index) over the lanes, ties to the smaller index
#pragma unroll
    for (int sh = 1; sh < NL; sh <<= 1) {
      const T v2 = lane_bcast(v, lane ^ sh);
      const int i2 = lane_bcast(ix, lane ^ sh);
      const bool take = (i2 < NB) && (ix == NB || v2 < v || (v2 == v && i2 < ix));
      v = take ? v2 : v;
      ix = take ? i2 : ix;
    }
    if (ix != l) {
      const T d_l = vget<NL, NW>(dv, l);
      const int p_l = vget<NL, NW>(pv, l);
      const int p_i = vget<NL, NW>(pv, ix);
      vset<NL, NW>(dv, kl, l, v);
      vset<NL, NW>(pv, kl, l, p_i);
      vset<NL, NW>(dv, kl, ix, d_l);
      vset<NL, NW>(pv, kl, ix, p_l);
    }
  }
  int *const posw = S.pos[wv];
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    const int idx = kl + q * NL;
    if (idx < n) {
      if (q == wv) wout[idx] = dv[q];
      posw[pv[q]] = idx;
    }
  }
  wsync();
  const int mcols = n - kopt_mdrop(solver);
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < n) {
      const int r = posw[j];
      if (k < n && r < mcols) A[k + (size_t)r * nm] = R[j];
    }
  }
}


template <typename T, int NB>
__global__ __launch_bounds__(64 * GroupShape<NB>::NW) void eig_wave_kernel(
    int L, int nm, int n, T *__restrict__ a, T *__restrict__ w, int *__restrict__ info, int solver,
    const int *__restrict__ only)
{
  constexpr int NL = GroupShape<NB>::NL, G = GroupShape<NB>::G;
  __shared__ WaveLds<T, NB> smem[G];
  const int lane = threadIdx.x & 63;
  const int wv = (int)(threadIdx.x >> 6);   // wave of the group (NW > 1)
  const int g = lane / NL, kl = lane % NL;
  const int k = wv * NL + kl;               // matrix row owned by this lane
  const long long mat = (long long)blockIdx.x * G + g;
  if (mat >= L) return;  // whole group leaves together (only for NW == 1; otherwise G == 1)
  if (only && only[mat] != ST_OVERFLOW) return;  // fallback pass of the split path
  wave_solve<T, NB>(a + (size_t)mat * (size_t)nm * (size_t)n, nm, n, w + (size_t)mat * n,
                    info ? info + mat : nullptr, solver, smem[g], lane, wv, kl, k);
}

} // namespace eigmi
