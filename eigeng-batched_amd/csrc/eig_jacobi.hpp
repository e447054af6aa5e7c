// eig_jacobi.hpp - two-sided cyclic Jacobi for n <= NB <= 64, one matrix per
// group of NB lanes, register resident (lane k owns row k of A and of Z).
//
// Ordering: the circle (round-robin) schedule on NB positions.  In every round
// position i is paired with position NB-1-i, so the pairs are compile-time
// register pairs; after the round, positions 1..NB-1 rotate by one (registers
// shift, rows move to the next lane), so NB-1 rounds meet every pair once (one
// sweep).  Positions n..NB-1 are zero padding: their pairs never rotate.
//
// Per round, each pair (p, q = NB-1-p) gets the rotation of the reference's
// jac_tiled (jac_tiled.hpp:79-95): t = 2a_pq / (b + sign(hypot(2a_pq, b), b)),
// b = a_qq - a_pp, c = 1/hypot(1, t), s = c t; then A <- J^T A J (columns
// lane-locally, rows by exchanging lanes p and q) and Z <- Z J.  A rotation is
// skipped when |a_pq| <= eps (|a_pp| + |a_qq|); the iteration stops after a
// sweep with no rotation or after `maxsw` sweeps (JAC_MAX_SWEEP for the solver
// entry points; info = -1 then), and nsweeps (if given) gets the count.
// The reference pivots one rotation at a time on the largest element
// (jac_tiled.hpp:33-64); the cyclic order converges to the same eigenpairs.
#pragma once
#include "common.hpp"
#include "eig_wave.hpp"

namespace eigmi {

constexpr int JAC_MAX_SWEEP = 40;

template <typename T, int NB>
struct alignas(16) JacLds {
  alignas(16) typename vec2<T>::type cs[NB / 2];  // (c, s) of pair (p, NB-1-p)
  alignas(16) int pos[NB];
};

template <typename T, int NB>
__global__ __launch_bounds__(64) void eig_jacobi_kernel(int L, int nm, int n, T *__restrict__ a,
                                                        T *__restrict__ w, int *__restrict__ info,
                                                        int opts = 0, int maxsw = JAC_MAX_SWEEP,
                                                        int *__restrict__ nsweeps = nullptr)
{
  constexpr int G = 64 / NB;
  constexpr int H = NB / 2;
  typedef typename vec2<T>::type T2;
  __shared__ JacLds<T, NB> smem[G];

  const int lane = threadIdx.x & 63;
  const int g = lane / NB, k = lane % NB;
  const int base = lane & ~(NB - 1);
  const long long mat = (long long)blockIdx.x * G + g;
  if (mat >= L) return;
  JacLds<T, NB> &S = smem[g];
  T *__restrict__ A = a + (size_t)mat * (size_t)nm * (size_t)n;

  // ---- load row k of the symmetric matrix (upper triangle only) -------------
  T RA[NB], RZ[NB];
  {
    const int kc = (k < n) ? k : n - 1;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int jc = (j < n) ? j : n - 1;
      const size_t idx = (kc <= jc) ? (size_t)kc + (size_t)jc * nm : (size_t)jc + (size_t)kc * nm;
      const T v = A[idx];
      RA[j] = (j < n && k < n) ? v : T(0);
      RZ[j] = (j == k) ? T(1) : T(0);
    }
  }
  int orig = k;   // original index of the element at position k (this lane's row)
  const T eps = eps_of<T>();

  int sweep = 0;
  for (; sweep < maxsw; ++sweep) {
    bool rotated = false;
    for (int round = 0; round < NB - 1; ++round) {
      // own diagonal a_kk and the pair element a_{k, NB-1-k}
      T dk = T(0), xk = T(0);
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        dk = (k == j) ? RA[j] : dk;
        xk = (k == j) ? RA[NB - 1 - j] : xk;
      }
      const int partner = base + (k ^ (NB - 1));
      const T dq = lane_bcast(dk, partner);
      // pair (p, q) with p < q; both lanes compute the same rotation
      const bool lo = k < H;
      const T app = lo ? dk : dq, aqq = lo ? dq : dk;
      // a_pq from row p only: rows p and q hold it twice, equal only up to rounding
      const T xq = lane_bcast(xk, partner);
      const T apq = lo ? xk : xq;
      T c = T(1), s = T(0);
      const bool rot = abs_(apq) > eps * (abs_(app) + abs_(aqq)) && apq != T(0);
      if (rot) {
        const T alpha = apq + apq, beta = aqq - app;
        const T t = alpha / (beta + sign_(hypot_c(alpha, beta), beta));
        c = T(1) / hypot1_c(t);
        s = c * t;
      }
      if (group_ballot<NB>(rot)) rotated = true;
      if (lo) S.cs[k] = T2{c, s};
      wsync();
      // A <- A J (columns p, q) and Z <- Z J, lane-local for every pair
#pragma unroll
      for (int p = 0; p < H; ++p) {
        const T2 q = S.cs[p];
        const T ap = RA[p], aq = RA[NB - 1 - p];
        RA[p] = q.x * ap - q.y * aq;
        RA[NB - 1 - p] = q.y * ap + q.x * aq;
        const T zp = RZ[p], zq = RZ[NB - 1 - p];
        RZ[p] = q.x * zp - q.y * zq;
        RZ[NB - 1 - p] = q.y * zp + q.x * zq;
      }
      // A <- J^T A: rows p and q mix (this lane is row k, its partner row NB-1-k)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const T other = lane_bcast(RA[j], partner);
        RA[j] = lo ? (c * RA[j] - s * other) : (s * other + c * RA[j]);
      }
      wsync();
      // circle move: positions 1..NB-1 shift by one (position NB-1 -> 1)
      {
        const int src = (k == 0) ? base : base + ((k == 1) ? NB - 1 : k - 1);
#pragma unroll
        for (int j = 0; j < NB; ++j) RA[j] = lane_bcast(RA[j], src);
        orig = lane_bcast(orig, src);
        const T la = RA[NB - 1], lz = RZ[NB - 1];
#pragma unroll
        for (int j = NB - 1; j >= 2; --j) { RA[j] = RA[j - 1]; RZ[j] = RZ[j - 1]; }
        RA[1] = la;
        RZ[1] = lz;
      }
    }
    if (!rotated) break;
  }

  // ---- eigenvalues at the positions, ascending rank, store --------------------
  T dk = T(0);
#pragma unroll
  for (int j = 0; j < NB; ++j) dk = (k == j) ? RA[j] : dk;
  if (info && k == 0) info[mat] = (sweep >= maxsw) ? -1 : 0;
  if (nsweeps && k == 0) nsweeps[mat] = sweep;
  // rank = number of real positions (orig < n) ordered before this one by
  // (value, original index); the reference's Jacobi leaves its output unsorted
  // (jac_tiled.hpp:143-147), the ABI returns ascending eigenvalues for every solver
  const bool real = orig < n;
  int rank = 0;
  for (int j = 0; j < NB; ++j) {
    const T dj = lane_bcast(dk, base + j);
    const int oj = lane_bcast(orig, base + j);
    const bool before = (oj < n) && (dj < dk || (dj == dk && oj < orig));
    rank += before ? 1 : 0;
  }
  if (kopt_unsorted(opts)) rank = orig;           // the rotations' order (the reference's)
  if (real) {
    w[(size_t)mat * n + rank] = dk;
    S.pos[k] = rank;
  }
  wsync();
  // column k of Z (register k) is the eigenvector of the element at position k;
  // lane r owns row r of Z = row r of the original basis
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int oj = lane_bcast(orig, base + j);
    if (oj < n) {
      const int r = S.pos[j];
      if (k < n && r < n - kopt_mdrop(opts)) A[k + (size_t)r * nm] = RZ[j];
    }
  }
}

} // namespace eigmi
