// eig_global.hpp - workgroup-per-matrix solver for 64 < n <= 512, matrices
// streamed from global memory (L2 / MALL): the working matrix lives in the
// caller's `a`, the eigenvector accumulator Z in a per-workgroup workspace
// slot (persistent grid, so the workspace is bounded by the grid size, not L).
//
// One workgroup of 256 threads per matrix; thread t owns rows t and t + 256:
//   1. symmetrise a in place (lower <- upper; only the upper triangle is read);
//   2. Householder tridiagonalisation, one column per step: p = beta A11 u and
//      the rank-2 update A11 += u v^T + v u^T are row-parallel sweeps over the
//      columns of A11 (coalesced), u and v are broadcast from LDS; u_c stays in
//      a(0:c-1, c), beta_c in LDS (tred1_tiled.hpp:42-212 convention);
//   3. Z = H_{n-1} ... H_1 by right-multiplication (row-parallel, coalesced);
//   4. QL: wave 0 runs the scalar recurrence of a sweep (d, e in LDS) and
//      stages its rotations in LDS; after a barrier every thread applies the
//      sweep to its rows of Z;
//   5. selection sort with the reference's tie rule, permuted copy Z -> a.
//
// Reference path replaced: eigen_GPU_batch_ (eigen_GPU_batch.cu:138-245) with
// hhsy2tr_ (hhsy2tr.hpp:3-326), tql2_/imtql2_ (tql2.hpp:3-214, imtql2.hpp:3-180)
// and hhtr2sy_ (hhtr2sy.hpp:3-220).
#pragma once
#include "common.hpp"
#include "eig_wave.hpp"

namespace eigmi {

constexpr int GL_THREADS = 256;  // default workgroup size (tunable: 256, 512, 1024)
constexpr int GL_NMAX = 512;

template <typename T>
struct alignas(16) GlobalLds {
  alignas(16) T u[GL_NMAX];
  alignas(16) T v[GL_NMAX];
  alignas(16) T d[GL_NMAX];
  alignas(16) T e[GL_NMAX];
  alignas(16) T beta[GL_NMAX];
  alignas(16) typename vec2<T>::type cs[GL_NMAX];
  alignas(16) int pos[GL_NMAX];
  alignas(16) T red[2][16];     // per-wave partial sums (up to 16 waves)
  int ctl[4];                    // wave 0 -> block: sweep bounds and status
};

template <int NT, typename T>
__device__ __forceinline__ T block_sum(T x, T (&red)[2][16], int slot)
{
  x = group_sum<64>(x);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[slot][wv] = x;
  __syncthreads();
  T s = T(0);
#pragma unroll
  for (int q = 0; q < NT / 64; ++q) s += red[slot][q];
  return s;
}

// Apply the staged rotations of one sweep (pairs l..m-1, from the top) to row r of Z.
template <typename T>
__device__ __forceinline__ void chain_row(T *__restrict__ Z, int n, int r, int l, int m,
                                          const typename vec2<T>::type *cs)
{
  T h1 = Z[r + (size_t)m * n];
  int i = m - 1;
  for (; i - 3 >= l; i -= 4) {
    const T z0 = Z[r + (size_t)i * n], z1 = Z[r + (size_t)(i - 1) * n];
    const T z2 = Z[r + (size_t)(i - 2) * n], z3 = Z[r + (size_t)(i - 3) * n];
    const typename vec2<T>::type q0 = cs[i], q1 = cs[i - 1], q2 = cs[i - 2], q3 = cs[i - 3];
    Z[r + (size_t)(i + 1) * n] = q0.y * z0 + q0.x * h1;
    h1 = q0.x * z0 - q0.y * h1;
    Z[r + (size_t)i * n] = q1.y * z1 + q1.x * h1;
    h1 = q1.x * z1 - q1.y * h1;
    Z[r + (size_t)(i - 1) * n] = q2.y * z2 + q2.x * h1;
    h1 = q2.x * z2 - q2.y * h1;
    Z[r + (size_t)(i - 2) * n] = q3.y * z3 + q3.x * h1;
    h1 = q3.x * z3 - q3.y * h1;
  }
  for (; i >= l; --i) {
    const T z0 = Z[r + (size_t)i * n];
    const typename vec2<T>::type q0 = cs[i];
    Z[r + (size_t)(i + 1) * n] = q0.y * z0 + q0.x * h1;
    h1 = q0.x * z0 - q0.y * h1;
  }
  Z[r + (size_t)l * n] = h1;
}

// Symmetrise A in place (lower <- upper) and reduce it to tridiagonal form by
// Householder reflections (tred1_tiled.hpp:42-212 convention): on return
// S.d / S.e hold the diagonal and off-diagonal (e(n-1) = 0), S.beta[c] the
// scale of reflector c, and u_c sits in A(0:c-1, c).
template <typename T, int NT>
__device__ __forceinline__ void global_tridiag(T *__restrict__ A, int nm, int n, GlobalLds<T> &S)
{
  const int t = threadIdx.x;
  // ---- 1. symmetrise in place ---------------------------------------------
  for (int c = 0; c < n; ++c)
    for (int r = c + 1 + t; r < n; r += NT) A[r + (size_t)c * nm] = A[c + (size_t)r * nm];
  __syncthreads();

  // ---- 2. tridiagonalisation ----------------------------------------------
  for (int c = n - 1; c >= 1; --c) {
    for (int r = t; r < c; r += NT) S.u[r] = A[r + (size_t)c * nm];
    if (t == 0) S.d[c] = A[c + (size_t)c * nm];
    __syncthreads();
    T part = T(0);
    for (int r = t; r < c; r += NT) part += S.u[r] * S.u[r];
    const T sig = block_sum<NT>(part, S.red, 0);
    const T ulast = S.u[c - 1];
    const T anorm = -sign_(sqrt_(sig), ulast);
    const T ul = ulast - anorm;
    const T beta = T(1) / nz(ul * anorm);
    __syncthreads();
    if (t == 0) {
      S.u[c - 1] = ul;
      S.e[c - 1] = anorm;
      S.beta[c] = beta;
      A[(c - 1) + (size_t)c * nm] = ul;
    }
    __syncthreads();
    T p0 = T(0), p1 = T(0);
    const int r0 = t, r1 = t + NT;
    if (r0 < c) {
      for (int j = 0; j < c; ++j) {
        const T uj = S.u[j];
        p0 = fma_(A[r0 + (size_t)j * nm], uj, p0);
        if (r1 < c) p1 = fma_(A[r1 + (size_t)j * nm], uj, p1);
      }
    }
    p0 = (r0 < c) ? p0 * beta : T(0);
    p1 = (r1 < c) ? p1 * beta : T(0);
    const T u0 = (r0 < c) ? S.u[r0] : T(0), u1 = (r1 < c) ? S.u[r1] : T(0);
    const T alpha = block_sum<NT>(p0 * u0 + p1 * u1, S.red, 1) * (beta * T(0.5));
    const T v0 = fma_(alpha, u0, p0), v1 = fma_(alpha, u1, p1);
    if (r0 < c) S.v[r0] = v0;
    if (r1 < c) S.v[r1] = v1;
    __syncthreads();
    if (r0 < c) {
      for (int j = 0; j < c; ++j) {
        const T uj = S.u[j], vj = S.v[j];
        T *p = A + r0 + (size_t)j * nm;
        *p = fma_(u0, vj, fma_(v0, uj, *p));
        if (r1 < c) {
          T *q = A + r1 + (size_t)j * nm;
          *q = fma_(u1, vj, fma_(v1, uj, *q));
        }
      }
    }
    __syncthreads();
  }
  if (t == 0) { S.d[0] = A[0]; S.e[n - 1] = T(0); }
}

// Whole solve of one matrix by the workgroup: A (leading dimension nm, order
// n), eigenvalue output wout[0..n-1], status *infop (may be null), Z an n x n
// accumulator private to the workgroup.
template <typename T, int NT>
__device__ __forceinline__ void global_solve(T *__restrict__ A, int nm, int n, T *__restrict__ wout,
                                             int *__restrict__ infop, int solver, T *__restrict__ Z,
                                             GlobalLds<T> &S)
{
  typedef typename vec2<T>::type T2;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const T tol = eps_of<T>() * T(sizeof(T) == 8 ? 512 : 16);

  global_tridiag<T, NT>(A, nm, n, S);

  // ---- 3. Z = H_{n-1} ... H_1 (right-multiplication, row-parallel) --------
  for (int j = 0; j < n; ++j)
    for (int r = t; r < n; r += NT) Z[r + (size_t)j * n] = (r == j) ? T(1) : T(0);
  __syncthreads();
  for (int c = n - 1; c >= 1; --c) {
    for (int r = t; r < c; r += NT) S.u[r] = A[r + (size_t)c * nm];
    __syncthreads();
    const T beta = S.beta[c];
    for (int r = t; r < n; r += NT) {
      T tr = T(0);
      for (int j = 0; j < c; ++j) tr = fma_(Z[r + (size_t)j * n], S.u[j], tr);
      tr *= beta;
      for (int j = 0; j < c; ++j) Z[r + (size_t)j * n] = fma_(tr, S.u[j], Z[r + (size_t)j * n]);
    }
    __syncthreads();
  }

  // ---- 4. QL: wave 0 drives, everyone applies ---------------------------
  for (int i = t; i < n; i += NT) S.cs[i] = T2{T(1), T(0)};
  __syncthreads();
  int err = 0;
  T shift = T(0), tst1 = T(0);   // explicit-shift state (wave 0)
  for (int l = 0; l < n && !err; ++l) {
    int itr = 0;
    bool tql_started = false;
    for (;;) {
      if (wv == 0) {
        int status = 0, m = l;                 // 0: converged, 1: sweep, 2: error
        if (kopt_solver(solver) == SOLVER_TQL2) {
          if (!tql_started) {
            const T h = (abs_(S.d[l]) + abs_(S.e[l])) * tol;
            tst1 = (h > tst1) ? h : tst1;
          }
          // first m >= l with |e(m)| <= tst1 (e(n-1) = 0 always qualifies)
          int mm = n - 1;
          for (int base = 0; base < n; base += 64) {
            const int idx = base + lane;
            const bool cand = idx >= l && idx < n && abs_(S.e[idx]) <= tst1;
            const uint64_t msk = __ballot(cand);
            if (msk) { mm = base + __builtin_ctzll(msk); break; }
          }
          m = tql_started ? m : mm;
          if (!tql_started) S.ctl[1] = mm;
          m = S.ctl[1];
          const bool done = tql_started ? (abs_(S.e[l]) <= tst1) : (m == l);
          if (done) status = 0;
          else if (itr >= QL_MAX_SWEEP) status = 2;
          else status = 1;
        } else {
          if (itr > QL_MAX_SWEEP) status = 2;
          else {
            int mm = n - 1;
            for (int base = 0; base < n; base += 64) {
              const int idx = base + lane;
              const bool cand = idx >= l && idx <= n - 2 &&
                                abs_(S.e[idx]) <= (abs_(S.d[idx]) + abs_(S.d[idx + 1])) * tol;
              const uint64_t msk = __ballot(cand);
              if (msk) { mm = base + __builtin_ctzll(msk); break; }
            }
            m = mm;
            status = (m == l) ? 0 : 1;
          }
        }
        if (status == 1) {
          if (kopt_solver(solver) == SOLVER_TQL2) {
            const T dl_old = S.d[l], el = S.e[l];
            const T p0 = (S.d[l + 1] - dl_old) / (el + el);
            const T r0 = hypot1_c(p0);
            const T psr = p0 + sign_(r0, p0);
            const T dl = el / psr, dl1 = el * psr;
            const T delta = dl_old - dl;
            wsync();
            for (int i = l + 2 + lane; i < n; i += 64) S.d[i] -= delta;
            if (lane == 0) { S.d[l] = dl; S.d[l + 1] = dl1; }
            shift += delta;
            wsync();
            T c = T(1), c2 = T(1), c3 = T(1), s = T(0), s2 = T(0);
            T pp = S.d[m];
            const T el1 = S.e[l + 1];
            T ei = S.e[m - 1], di = S.d[m - 1];
            for (int i = m - 1; i >= l; --i) {
              const int ip = (i > 0) ? i - 1 : 0;
              const T ein = S.e[ip], din = S.d[ip];
              c3 = c2; c2 = c; s2 = s;
              const T gg = c * ei, hh = c * pp;
              T r, cn, sn;
              if (!givens(ei, pp, r, cn, sn)) { cn = T(1); sn = T(0); }
              const T ei1 = s * r;
              s = sn;
              c = cn;
              pp = c * di - s * gg;
              const T di1 = hh + s * (c * gg + s * di);
              if (lane == 0) { S.e[i + 1] = ei1; S.d[i + 1] = di1; S.cs[i] = T2{c, s}; }
              ei = ein;
              di = din;
            }
            const T rr = (-s * s2 * c3 * el1 * S.e[l]) / dl1;
            wsync();
            if (lane == 0) { S.e[l] = s * rr; S.d[l] = c * rr; }
          } else {
            const T dl = S.d[l], el = S.e[l];
            const T f0 = (S.d[l + 1] - dl) / (el + el);
            const T g0 = hypot1_c(f0);
            T di1 = S.d[m];
            T r = (di1 - dl) + el / (f0 + sign_(g0, f0));
            T s = T(1), c = T(1), delta = T(0);
            T ei = S.e[m - 1], di = S.d[m - 1];
            int i;
            for (i = m - 1; i >= l; --i) {
              const int ip = (i > 0) ? i - 1 : 0;
              const T ein = S.e[ip], din = S.d[ip];
              const T f = s * ei, gg = r;
              const T bb = c * ei;
              if (!givens(f, gg, r, c, s)) break;
              const T dix = di1 - delta;
              di1 = di;
              const T q = (di1 - dix) * s + T(2) * c * bb;
              const T dx1 = fma_(s, q, dix);
              delta = dx1 - dix;
              if (lane == 0) { S.e[i + 1] = r; S.d[i + 1] = dx1; S.cs[i] = T2{c, s}; }
              r = c * q - bb;
              ei = ein;
              di = din;
            }
            const int j = (i + 1 > l) ? i + 1 : l;
            wsync();
            if (lane == 0) { S.d[j] -= delta; S.e[j] = r; S.e[m] = T(0); }
          }
        }
        if (lane == 0) { S.ctl[0] = status; S.ctl[1] = m; }
      }
      __syncthreads();
      const int status = S.ctl[0], m = S.ctl[1];
      if (status == 0) break;
      if (status == 2) { err = l + 1; break; }
      for (int r = t; r < n; r += NT) chain_row(Z, n, r, l, m, S.cs);
      __syncthreads();
      for (int i = l + t; i < m; i += NT) S.cs[i] = T2{T(1), T(0)};
      ++itr;
      tql_started = true;
      __syncthreads();
    }
    if (!err && kopt_solver(solver) == SOLVER_TQL2) {
      if (t == 0) S.d[l] += shift;
      __syncthreads();
    }
  }

  if (err) {
    for (int i = t; i < n; i += NT) wout[i] = S.d[i];
    if (infop && t == 0) *infop = err;
    __syncthreads();
    return;
  }
  if (infop && t == 0) *infop = 0;

  // ---- 5. sort (wave 0, reference tie rule) and permuted copy -------------
  if (wv == 0) {
    for (int i = lane; i < n; i += 64) S.pos[i] = i;
    wsync();
    const int nsort = kopt_unsorted(solver) ? 0 : n - 1;
    for (int l = 0; l < nsort; ++l) {
      T v = T(0);
      int ix = GL_NMAX;
      for (int idx = l + lane; idx < n; idx += 64) {
        const T di = S.d[idx];
        if (ix == GL_NMAX || di < v) { v = di; ix = idx; }
      }
#pragma unroll
      for (int sh = 1; sh < 64; sh <<= 1) {
        const T v2 = lane_bcast(v, lane ^ sh);
        const int i2 = lane_bcast(ix, lane ^ sh);
        const bool take = (i2 < GL_NMAX) && (ix == GL_NMAX || v2 < v || (v2 == v && i2 < ix));
        v = take ? v2 : v;
        ix = take ? i2 : ix;
      }
      if (ix != l) {
        const T d_l = S.d[l];
        const int p_l = S.pos[l], p_i = S.pos[ix];
        wsync();
        if (lane == 0) { S.d[l] = v; S.d[ix] = d_l; S.pos[l] = p_i; S.pos[ix] = p_l; }
        wsync();
      }
    }
  }
  __syncthreads();
  for (int i = t; i < n; i += NT) wout[i] = S.d[i];
  const int mcols = n - kopt_mdrop(solver);
  for (int j = 0; j < mcols; ++j) {
    const int src = S.pos[j];
    for (int r = t; r < n; r += NT) A[r + (size_t)j * nm] = Z[r + (size_t)src * n];
  }
  __syncthreads();
}

// `only` (may be null): solve just the matrices flagged ST_OVERFLOW there (the
// fallback pass of the split path).
template <typename T, int NT>
__global__ __launch_bounds__(NT) void eig_global_kernel(int L, int nm, int n, T *__restrict__ a,
                                                                T *__restrict__ w, int *__restrict__ info,
                                                                int solver, T *__restrict__ zws,
                                                                const int *__restrict__ only)
{
  __shared__ GlobalLds<T> S;
  T *__restrict__ Z = zws + (size_t)blockIdx.x * (size_t)n * (size_t)n;
  for (long long mat = blockIdx.x; mat < L; mat += gridDim.x)
    if (!only || only[mat] == ST_OVERFLOW)
      global_solve<T, NT>(a + (size_t)mat * (size_t)nm * (size_t)n, nm, n, w + (size_t)mat * n,
                    info ? info + mat : nullptr, solver, Z, S);
}

} // namespace eigmi
