// eig_split.hpp - n <= 64 solver split into three kernels so that the scalar QL
// recurrence runs one matrix per lane instead of one matrix per wave.
//
//   eig_tri_kernel    (group of NB lanes per matrix): load, Householder
//                     tridiagonalisation, Q = H_{n-1}...H_1 (wave_front), then
//                     Q -> workspace (n x n) and d, e -> workspace in a
//                     matrix-minor layout (d_i of matrix m at de[i*ld + m]).
//   eig_ql_lane_kernel (one matrix per lane, 64 per wave): IMTQL2 / TQL2 on the
//                     lane's d, e (LDS, element i of lane l at [i][l], bank
//                     conflict free), every Givens pair appended to the
//                     matrix's rotation log, every sweep to its sweep log as
//                     (l, m, lowest pair); then the reference's selection sort
//                     (tql2_tiled.hpp:149-176) gives w and the output slot of
//                     each column.
//   eig_apply_kernel  (group of NB lanes per matrix): Q back into registers,
//                     the logged sweeps replayed as rotation chains (lane-local,
//                     next sweep's log prefetched during the current chain),
//                     permuted store into a.
// A matrix whose log would exceed its capacity is flagged and solved again by
// the fused kernel (eig_wave_kernel with the `only` filter); its input is still
// untouched at that point because nothing above writes a.
#pragma once
#include "common.hpp"
#include "eig_wave.hpp"

namespace eigmi {

// rotation log capacity per matrix (pairs) and sweep log capacity; the
// measured maxima are 0.93 n^2 rotations / 1.7 n sweeps (FP32, n = 64) and
// 1.25 n^2 / 2.3 n (FP64, n = 32) over random and LETKF matrices
__host__ __device__ constexpr int rot_cap_of(int n) { return (3 * n * n) / 2 + 16; }
__host__ __device__ constexpr int sw_cap_of(int n) { return 3 * n + 8; }

// Indirection of the uneven-batch split route.  A class's matrices are the
// entries [begin, end) = ctl[0..1] of the sorted list (matrix indices, -1 in
// alignment gaps); chunk k takes the entries begin + chunk0 + idx, idx < cap,
// and solves entry idx in workspace slot idx (slots sized for order nmx).  The
// route runs only for a class of at least min_count entries (the fused list
// kernel takes smaller classes), which the kernels read on the device.
struct UvSplitArgs {
  const int *ctl;
  const int *list;
  const int *nv;
  const long long *aoff, *woff;
  int chunk0, cap, min_count, nmx;
};

__device__ __forceinline__ bool uv_split_entry(const UvSplitArgs &u, long long idx, int &mat)
{
  const int b = u.ctl[0], e = u.ctl[1];
  if (e - b < u.min_count || idx >= u.cap) return false;
  const long long pos = (long long)b + u.chunk0 + idx;
  if (pos >= e) return false;
  mat = u.list[pos];
  return mat >= 0;
}

// sweep code: l | m << 10 | lowest pair << 20  (n <= 1024)
__device__ __forceinline__ int sw_code(int l, int m, int ilo) { return l | (m << 10) | (ilo << 20); }
__device__ __forceinline__ int sw_m(int code) { return (code >> 10) & 1023; }
__device__ __forceinline__ int sw_ilo(int code) { return (code >> 20) & 1023; }

// LIST (NB = 64 and 128): the uneven-batch form (u), one workgroup per slot of
// the chunk, its own order n from nv, packed input at a + aoff, Q slot nmx^2
// (leading dimension n), d / e at rows i / nmx + i of de.
template <typename T, int NB, bool LIST = false>
__global__ __launch_bounds__(64 * GroupShape<NB>::NW) void eig_tri_kernel(int L, int nm, int n,
                                                                        const T *__restrict__ a,
                                                                        T *__restrict__ qws,
                                                                        T *__restrict__ de, int ld,
                                                                        UvSplitArgs u = UvSplitArgs{})
{
  // NB <= 64: groups of NB lanes; NB = 128: two waves per matrix, row k of
  // the matrix in lane k % 64 of wave k / 64
  constexpr int NL = GroupShape<NB>::NL, G = GroupShape<NB>::G, NW = GroupShape<NB>::NW;
  static_assert(!LIST || G == 1, "the list form takes one matrix per workgroup");
  __shared__ WaveLds<T, NB> smem[G];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane / NL, kl = lane % NL, k = kl + NL * wv;
  const long long mat = (long long)blockIdx.x * G + g;  // the slot (LIST) or the matrix
  const T *__restrict__ A;
  T *__restrict__ Q;
  int eoff = n;
  if constexpr (LIST) {
    int mi;
    if (!uv_split_entry(u, mat, mi)) return;         // the whole workgroup
    n = __builtin_amdgcn_readfirstlane(u.nv[mi]);
    nm = n;
    eoff = u.nmx;
    A = a + u.aoff[mi];
    Q = qws + (size_t)mat * u.nmx * u.nmx;
  } else {
    if (mat >= L) return;                            // the whole group (G = 1 when NW > 1)
    eoff = n;
    A = a + (size_t)mat * (size_t)nm * (size_t)n;
    Q = qws + (size_t)mat * (size_t)n * (size_t)n;
  }
  WaveLds<T, NB> &S = smem[g];
  T R[NB], dv[NW], ev[NW];
  wave_front<T, NB>(A, nm, n, k, kl, wv, S, R, dv, ev);
#pragma unroll
  for (int j = 0; j < NB; ++j)
    if (j < n && k < n) Q[k + (size_t)j * n] = R[j];
  if (wv == 0) {
    // d and e are lane-distributed (entry kl + NL h in dv[h]) and the same in every wave
#pragma unroll
    for (int h = 0; h < NW; ++h) {
      const int i = kl + NL * h;
      if (i < n) {
        de[(size_t)i * ld + mat] = dv[h];
        de[(size_t)(eoff + i) * ld + mat] = ev[h];
      }
    }
  }
}

// NLANE matrices per 64-lane workgroup (the rest of the lanes idle): the LDS
// holds 2 NMAX NLANE values, so large orders use fewer lanes per wave.
// LIST: the uneven-batch form (u): lane idx takes list entry idx of the chunk,
// its own order n from nv, d / e at rows i / nmx + i of de, eigenvalues at
// w + woff, output slots at slot + idx nmx; otherwise matrix idx of a uniform
// batch of order n.
template <typename T, int NMAX, int NLANE = 64, bool LIST = false>
__global__ __launch_bounds__(64) void eig_ql_lane_kernel(
    int L, int n, const T *__restrict__ de, int ld, typename vec2<T>::type *__restrict__ rotlog,
    int rot_cap, int *__restrict__ swlog, int sw_cap, int *__restrict__ nsw_out, int *__restrict__ status,
    int *__restrict__ slot, T *__restrict__ w, int *__restrict__ info, int solver, UvSplitArgs u = UvSplitArgs{})
{
  typedef typename vec2<T>::type T2;
  __shared__ T dS[NMAX][NLANE];
  __shared__ T eS[NMAX][NLANE];
  const int lane = threadIdx.x & 63;
  if (lane >= NLANE) return;
  const long long mat = (long long)blockIdx.x * NLANE + lane;   // the slot (LIST) or the matrix
  long long gmat = mat;                                         // the matrix (info)
  int eoff = n;                                                 // row of e(0) in de
  T *__restrict__ wout;
  int *__restrict__ sout;
  if constexpr (LIST) {
    int mi;
    if (!uv_split_entry(u, mat, mi)) return;
    gmat = mi;
    n = u.nv[mi];
    eoff = u.nmx;
    wout = w + u.woff[mi];
    sout = slot + (size_t)mat * u.nmx;
  } else {
    if (mat >= L) return;
    wout = w + (size_t)mat * n;
    sout = slot + (size_t)mat * n;
  }
  for (int i = 0; i < n; ++i) {
    dS[i][lane] = de[(size_t)i * ld + mat];
    eS[i][lane] = de[(size_t)(eoff + i) * ld + mat];
  }
  T2 *__restrict__ lg = rotlog + (size_t)mat * rot_cap;
  int *__restrict__ sw = swlog + (size_t)mat * sw_cap;
  const T tol = eps_of<T>() * T(sizeof(T) == 8 ? 512 : 16);
  int nrot = 0, nsw = 0, st = 0;

  if (kopt_solver(solver) == SOLVER_TQL2) {
    // explicit shift (tql2_tiled.hpp:40-146), as one flat loop of sweeps per
    // lane like the implicit-shift form below: l is set up (tst1, m; the
    // converged ones finished) when its predecessor converges, every pass
    // runs one sweep.  Each lane's arithmetic is the nested form's.
    T shift = T(0), tst1 = T(0);
    int l = 0, m = 0, itr = 0;
    bool fresh = true;
    while (l < n) {
      if (fresh) {
        for (;;) {
          const T h = (abs_(dS[l][lane]) + abs_(eS[l][lane])) * tol;
          tst1 = (h > tst1) ? h : tst1;
          m = l;
          while (m < n - 1 && abs_(eS[m][lane]) > tst1) ++m;
          if (m != l) break;
          dS[l][lane] += shift;
          if (++l >= n) break;
        }
        if (l >= n) break;
        fresh = false;
        itr = 0;
      }
      if (itr >= QL_MAX_SWEEP) { st = l + 1; break; }
      if (nsw >= sw_cap || nrot + (m - l) > rot_cap) { st = ST_OVERFLOW; break; }
      const T dl_old = dS[l][lane], el = eS[l][lane];
      const T p0 = (dS[l + 1][lane] - dl_old) / (el + el);
      const T r0 = hypot1_c(p0);
      const T psr = p0 + sign_(r0, p0);
      const T dl = el / psr, dl1 = el * psr;
      const T delta = dl_old - dl;
      dS[l][lane] = dl;
      dS[l + 1][lane] = dl1;
      for (int i = l + 2; i < n; ++i) dS[i][lane] -= delta;
      shift += delta;
      T c = T(1), c2 = T(1), c3 = T(1), s = T(0), s2 = T(0);
      T pp = dS[m][lane];
      const T el1 = eS[l + 1][lane];
      for (int i = m - 1; i >= l; --i) {
        c3 = c2; c2 = c; s2 = s;
        const T ei = eS[i][lane], di = dS[i][lane];
        const T gg = c * ei, hh = c * pp;
        T r, cn, sn;
        if (!givens(ei, pp, r, cn, sn)) { cn = T(1); sn = T(0); }
        eS[i + 1][lane] = s * r;
        s = sn;
        c = cn;
        pp = c * di - s * gg;
        dS[i + 1][lane] = hh + s * (c * gg + s * di);
        lg[nrot++] = T2{c, s};
      }
      const T rr = (-s * s2 * c3 * el1 * el) / dl1;
      eS[l][lane] = s * rr;
      dS[l][lane] = c * rr;
      sw[nsw++] = sw_code(l, m, l);
      ++itr;
      if (abs_(eS[l][lane]) <= tst1) {             // eigenvalue l converged
        dS[l][lane] += shift;
        ++l;
        fresh = true;
      }
    }
  } else {
    // implicit shift (imtql2_tiled.hpp:43-134) as one flat loop: every pass
    // runs one sweep of the lane's current eigenvalue l, moving l on past the
    // converged ones first.  Nested loops over l and its sweeps would make the
    // lanes of a wave wait for each other at every l (a wave's lanes leave a
    // loop together), costing the maximum over lanes of the sweeps of each l
    // instead of the maximum of their sums.  Each lane's sequence of sweeps,
    // and so its arithmetic, is the nested form's.
    int l = 0, itr = 0;
    while (l < n) {
      int m = l;
      for (;;) {
        if (itr > QL_MAX_SWEEP) break;
        m = l;
        while (m < n - 1 && abs_(eS[m][lane]) > (abs_(dS[m][lane]) + abs_(dS[m + 1][lane])) * tol) ++m;
        if (m != l) break;
        itr = 0;
        if (++l >= n) break;
      }
      if (l >= n) break;                           // every eigenvalue converged
      if (itr > QL_MAX_SWEEP) { st = l + 1; break; }
      ++itr;
      if (nsw >= sw_cap || nrot + (m - l) > rot_cap) { st = ST_OVERFLOW; break; }
      const T dl = dS[l][lane], el = eS[l][lane];
      const T f0 = (dS[l + 1][lane] - dl) / (el + el);
      const T g0 = hypot1_c(f0);
      T di1 = dS[m][lane];
      T r = (di1 - dl) + el / (f0 + sign_(g0, f0));
      T s = T(1), c = T(1), delta = T(0);
      T ei = eS[m - 1][lane], di = dS[m - 1][lane];
      T2 *lp = lg + nrot;
      int j = l;                                   // where the sweep ends
      for (int i = m - 1; i >= l; --i) {
        const int ip = (i > 0) ? i - 1 : 0;
        const T ein = eS[ip][lane], din = dS[ip][lane];
        const T f = s * ei, gg = r;
        const T bb = c * ei;
        const T ss = fma_(f, f, gg * gg);
        const bool brk = ss == T(0);
        const bool any_brk = __ballot(brk) != 0ull;
        // f = g = 0 (the underflow exit) runs as the rotation c = s = 0, which
        // stores d(i+1) - p; the exit's other effects follow below
        givens_nz(f, gg, brk ? std::numeric_limits<T>::min() : ss, r, c, s);
        const T dix = di1 - delta;
        di1 = di;
        const T q = (di1 - dix) * s + T(2) * c * bb;
        const T dx1 = fma_(s, q, dix);
        delta = dx1 - dix;
        eS[i + 1][lane] = r;
        dS[i + 1][lane] = dx1;
        *lp++ = T2{c, s};
        r = c * q - bb;
        ei = ein;
        di = din;
        if (__builtin_expect(any_brk, 0)) {
          // the exit: no rotation logged, e(i+1) = 0 (stored below), p
          // already taken off d(i+1), and this lane's sweep ends here
          if (brk) {
            --lp;
            r = T(0);
            delta = T(0);
            j = i + 1;
            i = l;
          }
        }
      }
      nrot = (int)(lp - lg);
      dS[j][lane] -= delta;
      eS[j][lane] = r;
      eS[m][lane] = T(0);
      sw[nsw++] = sw_code(l, m, j);
    }
  }

  status[mat] = st;
  nsw_out[mat] = nsw;
  if (st == ST_OVERFLOW) return;                 // solved again by the fused kernel
  if (info) info[gmat] = st;
  if (st != 0) {
    for (int i = 0; i < n; ++i) wout[i] = dS[i][lane];
    return;
  }
  // selection sort with the reference's rule; eS now holds the positions
  for (int i = 0; i < n; ++i) eS[i][lane] = T(i);
  const int nsort = kopt_unsorted(solver) ? 0 : n - 1;
  for (int l = 0; l < nsort; ++l) {
    T dl = dS[l][lane];
    int il = l;
    for (int j = l + 1; j < n; ++j) {
      const T dj = dS[j][lane];
      if (dl > dj) { dl = dj; il = j; }
    }
    if (il != l) {
      const T pl = eS[l][lane];
      eS[l][lane] = eS[il][lane];
      eS[il][lane] = pl;
      dS[il][lane] = dS[l][lane];
      dS[l][lane] = dl;
    }
  }
  // output position of each column; the columns beyond m are not stored (-1)
  const int mcols = n - kopt_mdrop(solver);
  for (int i = 0; i < n; ++i) {
    wout[i] = dS[i][lane];
    sout[(int)eS[i][lane]] = i < mcols ? i : -1;
  }
}

// Bit mask over the off-diagonal indices 0 .. NB-2 of the register QL kernel:
// one 64-bit word up to NB = 64, two beyond.
template <int NB>
struct QlMask {
  static constexpr int W = NB <= 64 ? 1 : 2;
  uint64_t w[W];
};
// bits l, l+1, ... set
template <int NB>
__device__ __forceinline__ QlMask<NB> qm_from(int l)
{
  QlMask<NB> m;
  m.w[0] = (l < 64) ? (~0ull << (l < 0 ? 0 : l)) : 0ull;
  if constexpr (QlMask<NB>::W == 2) m.w[1] = (l <= 64) ? ~0ull : ((l < 128) ? (~0ull << (l - 64)) : 0ull);
  return m;
}
template <int NB>
__device__ __forceinline__ QlMask<NB> qm_and(const QlMask<NB> &a, const QlMask<NB> &b)
{
  QlMask<NB> m;
#pragma unroll
  for (int q = 0; q < QlMask<NB>::W; ++q) m.w[q] = a.w[q] & b.w[q];
  return m;
}
template <int NB>
__device__ __forceinline__ QlMask<NB> qm_andnot(const QlMask<NB> &a, const QlMask<NB> &b)   // a & ~b
{
  QlMask<NB> m;
#pragma unroll
  for (int q = 0; q < QlMask<NB>::W; ++q) m.w[q] = a.w[q] & ~b.w[q];
  return m;
}
template <int NB>
__device__ __forceinline__ bool qm_zero(const QlMask<NB> &a)
{
  if constexpr (QlMask<NB>::W == 2) return (a.w[0] | a.w[1]) == 0ull;
  else return a.w[0] == 0ull;
}
// lowest set bit (the mask is not empty)
template <int NB>
__device__ __forceinline__ int qm_ctz(const QlMask<NB> &a)
{
  if constexpr (QlMask<NB>::W == 2) return a.w[0] ? __builtin_ctzll(a.w[0]) : 64 + __builtin_ctzll(a.w[1]);
  else return __builtin_ctzll(a.w[0]);
}

// IMTQL2 recurrence, one matrix per lane, with d and e in registers.
//
// All lanes of the wave walk the same register index at the same time: each
// round every lane finds its window [l, m] from a bit mask of negligible
// off-diagonals, then the chase over i = NB-2 .. 0 is unrolled and a lane acts
// only on the steps inside its own window.  No array is indexed at run time,
// so d and e stay in VGPRs.  The per-step arithmetic, the deflation test, the
// sweep limit and the logs are those of eig_ql_lane_kernel's IMTQL2 branch
// (imtql2_tiled.hpp:43-164), and the final ordering is the reference's
// selection sort (tql2_tiled.hpp:149-176), also unrolled over registers.
// LIST: the uneven-batch form (u): lane slot idx holds list entry idx of the
// chunk, with its own order (d / e at rows i / nmx + i of de), as in
// eig_ql_lane_kernel.
template <typename T, int NB, bool LIST = false>
__global__ __launch_bounds__(64) void eig_ql_reg_kernel(
    int L, int n, const T *__restrict__ de, int ld, typename vec2<T>::type *__restrict__ rotlog,
    int rot_cap, int *__restrict__ swlog, int sw_cap, int *__restrict__ nsw_out, int *__restrict__ status,
    int *__restrict__ slot, T *__restrict__ w, int *__restrict__ info, int opts = 0, UvSplitArgs u = UvSplitArgs{})
{
  typedef typename vec2<T>::type T2;
  static_assert(NB <= 128, "at most two 64-bit mask words per lane");
  const int lane = threadIdx.x & 63;
  const long long mat = (long long)blockIdx.x * 64 + lane;   // the slot (LIST) or the matrix
  bool live = mat < L;
  long long gmat = mat;                                        // the matrix (info)
  int eoff = n;                                                // row of e(0) in de
  T *__restrict__ wout;
  int *__restrict__ sout;
  if constexpr (LIST) {
    int mi = 0;
    live = live && uv_split_entry(u, mat, mi);
    gmat = mi;
    n = live ? u.nv[mi] : 1;                                   // this lane's order
    eoff = u.nmx;
    wout = w + (live ? u.woff[mi] : 0);
    sout = slot + (size_t)(live ? mat : 0) * u.nmx;
  } else {
    wout = w + (size_t)(live ? mat : 0) * n;
    sout = slot + (size_t)(live ? mat : 0) * n;
  }
  const long long mt = live ? mat : 0;
  T d[NB], e[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    d[i] = (live && i < n) ? de[(size_t)i * ld + mt] : T(0);
    e[i] = (live && i < n) ? de[(size_t)(eoff + i) * ld + mt] : T(0);
  }
  T2 *__restrict__ lg = rotlog + (size_t)mt * rot_cap;
  int *__restrict__ sw = swlog + (size_t)mt * sw_cap;
  const T tol = eps_of<T>() * T(sizeof(T) == 8 ? 512 : 16);
  // bits from n-1 on always set: e(n-1) = 0 ends every window at n-1 at the latest
  const QlMask<NB> top = qm_from<NB>(n - 1);
  int l = 0, itr = 0, st = 0, nrot = 0, nsw = 0;
  bool run = live && n > 1;

  int lmin = 0;  // wave minimum of l over running lanes (l never decreases)
  while (__ballot(run)) {
    // ---- window of this round ----------------------------------------------
    // negligible off-diagonals below every running lane's l are never looked
    // at again, so 8-blocks entirely below lmin are skipped
    QlMask<NB> small = top;
#pragma unroll
    for (int b = 0; b < NB / 8; ++b) {
      if (8 * b + 8 > lmin) {
#pragma unroll
        for (int i = 8 * b; i < 8 * b + 8 && i < NB - 1; ++i) {
          // NaN counts as negligible; the bits from n-1 on are set in top already
          const bool sm = !(abs_(e[i]) > (abs_(d[i]) + abs_(d[i + 1])) * tol);
          small.w[i >> 6] |= sm ? (1ull << (i & 63)) : 0ull;
        }
      }
    }
    if (run && itr > QL_MAX_SWEEP) { st = l + 1; run = false; }
    const QlMask<NB> cand = qm_andnot(qm_from<NB>(l), small);
    if (run && qm_zero(cand)) run = false;       // every eigenvalue converged
    int m = 0;
    if (run) {
      const int nl = qm_ctz(cand);
      if (nl != l) itr = 0;
      l = nl;
      m = qm_ctz(qm_and(small, qm_from<NB>(l)));
      if (nsw >= sw_cap || nrot + (m - l) > rot_cap) { st = ST_OVERFLOW; run = false; }
    }
    const bool sweep = run;
    // this round's pairs: lo <= i <= hi (empty for idle lanes); wave union
    int lo = sweep ? l : NB + 8, hi = sweep ? m - 1 : -1;
    int lo_w = lo, hi_w = hi;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const int lo2 = __shfl_xor(lo_w, off), hi2 = __shfl_xor(hi_w, off);
      lo_w = lo2 < lo_w ? lo2 : lo_w;
      hi_w = hi2 > hi_w ? hi2 : hi_w;
    }
    lo_w = __builtin_amdgcn_readfirstlane(lo_w);
    hi_w = __builtin_amdgcn_readfirstlane(hi_w);
    lmin = lo_w;
    // d(l), e(l), d(l+1), d(m) of each lane (per-lane indices: select chain)
    T dl = T(0), el = T(1), dl1 = T(0), dm = T(0);
#pragma unroll
    for (int b = 0; b < NB / 8; ++b) {
      if (8 * b + 8 > lo_w && 8 * b <= hi_w + 1) {
#pragma unroll
        for (int i = 8 * b; i < 8 * b + 8; ++i) {
          if (i < NB - 1) {
            const bool at = (i == lo);
            dl = at ? d[i] : dl;
            el = at ? e[i] : el;
            dl1 = at ? d[i + 1] : dl1;
          }
          dm = (i == hi + 1) ? d[i] : dm;
        }
      }
    }
    T r = T(0);
    if (sweep) {
      const T f0 = (dl1 - dl) / (el + el);
      const T g0 = hypot1_c(f0);
      r = (dm - dl) + el / (f0 + sign_(g0, f0));
    }
    // ---- chase, i = m-1 down to l (lockstep over the wave's union) ----------
    // the window bounds go through an empty asm so that the step predicates
    // are formed where they are used instead of being kept from above
    asm volatile("" : "+v"(lo), "+v"(hi));
    const unsigned span = sweep ? (unsigned)(hi - lo) : 0u;   // idle lanes: i - lo < 0 fails
    T s = T(1), c = T(1), delta = T(0);
    int ilo = l;
    T2 *lp = lg + nrot;   // next free slot of this lane's rotation log
#pragma unroll
    for (int b = NB / 8 - 1; b >= 0; --b) {
      if (8 * b <= hi_w && 8 * b + 7 >= lo_w) {
#pragma unroll
        for (int i = 8 * b + 7; i >= 8 * b; --i) {
          if (i <= NB - 2) {
            // only the lanes whose window holds i run the step (the exec
            // mask), so the values carried along the chase need no selects.
            // The step is EISPACK's: s = f/r, c = g/r, g = d(i+1) - p,
            // r = (d(i) - g) s + 2 c b, p = s r, d(i+1) = g + p, g = c r - b;
            // e(i+1) = r except e(m) = 0; at the end of the window d(l) -= p,
            // e(l) = g.
            if ((unsigned)(i - lo) <= span) {
              const T ei = e[i];
              const T f = s * ei, bb = c * ei;
              const T ss = fma_(f, f, r * r);
              const bool brk = ss == T(0);
              const bool any_brk = __ballot(brk) != 0ull;
              // f = g = 0 (the underflow exit) runs as the rotation c = s = 0,
              // which leaves d(i+1) - p; the exit's other effects follow below
              T rn, cn, sn;
              givens_nz(f, r, brk ? std::numeric_limits<T>::min() : ss, rn, cn, sn);
              const T dix = d[i + 1] - delta;
              const T q = (d[i] - dix) * sn + T(2) * cn * bb;
              const T dx1 = fma_(sn, q, dix);
              *lp++ = T2{cn, sn};
              e[i + 1] = (i == hi) ? T(0) : rn;
              d[i + 1] = dx1;
              delta = dx1 - dix;
              r = cn * q - bb;
              c = cn;
              s = sn;
              d[i] = (i == lo) ? d[i] - delta : d[i];
              e[i] = (i == lo) ? r : e[i];
              if (__builtin_expect(any_brk, 0)) {
                // the exit: no rotation logged, e(i+1) = 0, e(l) untouched,
                // and the lane's sweep ends here
                if (brk) {
                  --lp;
                  e[i + 1] = T(0);
                  e[i] = ei;
                  ilo = i + 1;
                  lo = NB + 8;
                }
              }
            }
          }
        }
      }
    }
    nrot = (int)(lp - lg);
    if (sweep) {
      sw[nsw++] = sw_code(l, m, ilo);
      ++itr;
    }
  }

  if (!live) return;
  status[mat] = st;
  nsw_out[mat] = nsw;
  if (st == ST_OVERFLOW) return;                 // solved again by the fused kernel
  if (info) info[gmat] = st;
  if (st != 0) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i < n) wout[i] = d[i];
    return;
  }
  // selection sort with the reference's rule; e now holds the positions
#pragma unroll
  for (int i = 0; i < NB; ++i) e[i] = T(i);
  const int nsort = kopt_unsorted(opts) ? 0 : n - 1;
#pragma unroll
  for (int l2 = 0; l2 < NB - 1; ++l2) {
    if (l2 < nsort) {
      T dm = d[l2], pm = e[l2];
      int im = l2;
#pragma unroll
      for (int j = l2 + 1; j < NB; ++j) {
        const bool lt = (j < n) && (dm > d[j]);
        dm = lt ? d[j] : dm;
        pm = lt ? e[j] : pm;
        im = lt ? j : im;
      }
      const T dl2 = d[l2], pl2 = e[l2];
#pragma unroll
      for (int j = l2 + 1; j < NB; ++j) {
        const bool hit = (j == im);
        d[j] = hit ? dl2 : d[j];
        e[j] = hit ? pl2 : e[j];
      }
      d[l2] = dm;
      e[l2] = pm;
    }
  }
  const int mcols = n - kopt_mdrop(opts);
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (i < n) {
      wout[i] = d[i];
      sout[(int)e[i]] = i < mcols ? i : -1;
    }
  }
}

template <typename T, int NB>
struct alignas(16) ApplyLds {
  // sweeps in flight: two for FP32 NB = 128 (apply_two_sweeps; it measured
  // slower in FP64 and no faster at NB = 64), one otherwise
  static constexpr int NS = (sizeof(T) == 4 && NB == 128) ? 2 : 1;
  alignas(16) typename vec2<T>::type cs[NS][NB];   // rotations of the sweeps in flight
  alignas(16) int pos[NB];
};

// NB <= 64: groups of NB lanes, lane k = row k of Q.  NB = 128: one wave per
// 64-row block of a matrix (NR = 2 workgroups per matrix), each lane a full
// row of 128 columns; the blocks share nothing, every wave stages the
// rotations of a sweep (up to NB - 1) itself, NR per lane.
// LIST (NB = 64 and 128): the uneven-batch form (u), slot = blockIdx / NR.
// FP32 NB = 128 replays the sweeps in pairs, the second sweep one 8-block
// behind the first (apply_two_sweeps), so that two rotation chains overlap.
// occupancy floor of the apply kernel (waves per SIMD) with two sweeps in
// flight: the row in registers plus two sweeps' rotation blocks
template <typename T, int NB> struct ApplyWaves {
  static constexpr int value = (ApplyLds<T, NB>::NS == 2) ? 3 : 1;
};

template <typename T, int NB, bool LIST = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(ApplyWaves<T, NB>::value))) void eig_apply_kernel(int L, int nm, int n, T *__restrict__ a,
                                                       const T *__restrict__ qws,
                                                       const typename vec2<T>::type *__restrict__ rotlog,
                                                       int rot_cap, const int *__restrict__ swlog, int sw_cap,
                                                       const int *__restrict__ nsw_in,
                                                       const int *__restrict__ status, const int *__restrict__ slot,
                                                       UvSplitArgs u = UvSplitArgs{})
{
  typedef typename vec2<T>::type T2;
  constexpr int NL = GroupShape<NB>::NL, G = GroupShape<NB>::G, NR = NB / NL;
  static_assert(!LIST || G == 1, "the list form takes one matrix per NR workgroups");
  __shared__ ApplyLds<T, NB> smem[G];
  const int lane = threadIdx.x & 63;
  const int g = lane / NL, kl = lane % NL;
  const long long mat = (NR == 1) ? (long long)blockIdx.x * G + g : (long long)(blockIdx.x / NR);
  const int k = (NR == 1) ? kl : kl + NL * (int)(blockIdx.x % NR);   // row of Q
  T *__restrict__ A;
  const T *__restrict__ Q;
  size_t sstride = n;
  if constexpr (LIST) {
    int mi;
    if (!uv_split_entry(u, mat, mi)) return;
    n = __builtin_amdgcn_readfirstlane(u.nv[mi]);
    nm = n;
    if (k - kl >= n) return;                         // a row block beyond the order
    A = a + u.aoff[mi];
    Q = qws + (size_t)mat * u.nmx * u.nmx;
    sstride = u.nmx;
  } else {
    if (mat >= L) return;
    A = a + (size_t)mat * (size_t)nm * (size_t)n;
    Q = qws + (size_t)mat * (size_t)n * (size_t)n;
  }
  if (status[mat] != 0) return;
  ApplyLds<T, NB> &S = smem[g];
  const T2 *__restrict__ lg = rotlog + (size_t)mat * rot_cap;
  const int *__restrict__ sw = swlog + (size_t)mat * sw_cap;

  T R[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) R[j] = (j < n && k < n) ? Q[k + (size_t)j * n] : T(0);
  // the row's loads complete here, after all of them were issued: left
  // pending into the sweep loop, the wait for them inside the loop would also
  // wait for the next sweep's prefetched rotations, on every sweep
#pragma unroll
  for (int j = 0; j < NB; ++j) asm volatile("" : "+v"(R[j]));
  if constexpr (ApplyLds<T, NB>::NS == 2) {
#pragma unroll
    for (int t = 0; t < NR; ++t) {
      S.cs[0][kl + NL * t] = T2{T(1), T(0)};
      S.cs[1][kl + NL * t] = T2{T(1), T(0)};
    }
    const int nsw = nsw_in[mat];
    // sweeps in pairs (the second one block behind the first, apply_two_sweeps);
    // software pipeline: the next pair's headers and this lane's log entries
    // (positions kl + NL t of each sweep) are loaded while the current pair runs
    int code0 = (nsw > 0) ? sw[0] : 0, code1 = (nsw > 1) ? sw[1] : 0;
    int base = 0;
    T2 q0[NR], q1[NR];
    {
      const int cnt0 = sw_m(code0) - sw_ilo(code0), cnt1 = sw_m(code1) - sw_ilo(code1);
#pragma unroll
      for (int t = 0; t < NR; ++t) {
        q0[t] = (nsw > 0 && kl + NL * t < cnt0) ? lg[kl + NL * t] : T2{T(1), T(0)};
        q1[t] = (nsw > 1 && kl + NL * t < cnt1) ? lg[cnt0 + kl + NL * t] : T2{T(1), T(0)};
      }
    }
    for (int s = 0; s < nsw; s += 2) {
      const bool two = s + 1 < nsw;
      const int m0 = sw_m(code0), ilo0 = sw_ilo(code0), cnt0 = m0 - ilo0;
      const int m1 = sw_m(code1), ilo1 = sw_ilo(code1), cnt1 = two ? m1 - ilo1 : 0;
#pragma unroll
      for (int t = 0; t < NR; ++t) {
        if (kl + NL * t < cnt0) S.cs[0][m0 - 1 - (kl + NL * t)] = q0[t];
        if (kl + NL * t < cnt1) S.cs[1][m1 - 1 - (kl + NL * t)] = q1[t];
      }
      base += cnt0 + cnt1;
      // prefetch sweeps s+2, s+3
      const int c2 = (s + 2 < nsw) ? sw[s + 2] : 0, c3 = (s + 3 < nsw) ? sw[s + 3] : 0;
      {
        const int n2 = (s + 2 < nsw) ? sw_m(c2) - sw_ilo(c2) : 0;
        const int n3 = (s + 3 < nsw) ? sw_m(c3) - sw_ilo(c3) : 0;
#pragma unroll
        for (int t = 0; t < NR; ++t) {
          if (kl + NL * t < n2) q0[t] = lg[base + kl + NL * t];
          if (kl + NL * t < n3) q1[t] = lg[base + n2 + kl + NL * t];
        }
      }
      wsync();
      // an odd last sweep goes with an empty second window (its blocks are
      // no-ops: S.cs[1] holds identities only); one code path keeps the row in
      // the same registers
      apply_two_sweeps<T, NB>(R, S.cs[0], ilo0, m0 - 1, S.cs[1], two ? ilo1 : NB + 8, two ? m1 - 1 : -1);
      wsync();
#pragma unroll
      for (int t = 0; t < NR; ++t) {
        const int i = kl + NL * t;
        if (i >= ilo0 && i <= m0 - 1) S.cs[0][i] = T2{T(1), T(0)};
        if (two && i >= ilo1 && i <= m1 - 1) S.cs[1][i] = T2{T(1), T(0)};
      }
      code0 = c2;
      code1 = c3;
      wsync();
    }
  } else {
#pragma unroll
    for (int t = 0; t < NR; ++t) S.cs[0][kl + NL * t] = T2{T(1), T(0)};
    const int nsw = nsw_in[mat];
    // software pipeline: the next sweep's header and this lane's log entries
    // (positions kl + NL t of the sweep) are loaded while the current chain runs
    int code = (nsw > 0) ? sw[0] : 0;
    int base = 0;
    T2 q[NR];
    {
      const int m = sw_m(code), ilo = sw_ilo(code);
#pragma unroll
      for (int t = 0; t < NR; ++t) q[t] = (nsw > 0 && kl + NL * t < m - ilo) ? lg[kl + NL * t] : T2{T(1), T(0)};
    }
    for (int s = 0; s < nsw; ++s) {
      const int m = sw_m(code), ilo = sw_ilo(code);
      const int cnt = m - ilo;
#pragma unroll
      for (int t = 0; t < NR; ++t)
        if (kl + NL * t < cnt) S.cs[0][m - 1 - (kl + NL * t)] = q[t];
      base += cnt;
      // prefetch sweep s+1
      const int code2 = (s + 1 < nsw) ? sw[s + 1] : 0;
      {
        const int m2 = sw_m(code2), ilo2 = sw_ilo(code2);
#pragma unroll
        for (int t = 0; t < NR; ++t)
          if (s + 1 < nsw && kl + NL * t < m2 - ilo2) q[t] = lg[base + kl + NL * t];
      }
      wsync();
      apply_sweep_rotations<T, NB>(R, S.cs[0], ilo, m - 1);
      wsync();
#pragma unroll
      for (int t = 0; t < NR; ++t) {
        const int i = kl + NL * t;
        if (i >= ilo && i <= m - 1) S.cs[0][i] = T2{T(1), T(0)};
      }
      code = code2;
      wsync();
    }
  }
#pragma unroll
  for (int t = 0; t < NR; ++t)
    if (kl + NL * t < n) S.pos[kl + NL * t] = slot[(size_t)mat * sstride + kl + NL * t];
  wsync();
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < n) {
      const int r = S.pos[j];
      if (k < n && r >= 0) A[k + (size_t)r * nm] = R[j];
    }
  }
}

} // namespace eigmi
