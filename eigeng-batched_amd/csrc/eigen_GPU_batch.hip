// eigen_GPU_batch.hip - host dispatch and C ABI of the gfx950 batched
// symmetric eigensolver (libeigenGbatch).
//
// Replaces the reference's host layer: eigen_GPU_batch_RUN
// (eigen_GPU_batch.cu:248-325), the C symbols eigen_GPU_batch_DP/_FP
// (eigen_GPU_batch.cu:327-359) and the launch heuristics of
// eigen_GPU_batch_get_Launch_params (eigen_GPU_batch.hpp:23-102).
// No device query is made per call: the launch shape depends only on n and L.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "../../include/eigen_GPU_batch.h"
#include "eig_wave.hpp"
#include "eig_global.hpp"
#include "eig_jacobi.hpp"
#include "eig_split.hpp"
#include "eig_uneven.hpp"
#include "eig_split_large.hpp"
#include "eig_tri_reg.hpp"
#include "eig_check.hpp"
#include "eig_symdmt.hpp"
#include "eig_gemm.hpp"
#include "eig_refine.hpp"
#include "eig_layout.hpp"
#include "eig_warm.hpp"
#include "eig_warmnet.hpp"
#include "eig_wy.hpp"

#include <stdio.h>
#include <vector>

#include <mutex>
#include <map>
#include <utility>

using namespace eigmi;

namespace {

bool bad_args(int L, int nm, int n, int m) { return L < 1 || nm < n || n < 1 || m < 1 || m > n; }

constexpr int WS_HEAD = 256;      // bytes; keeps the workspace non-empty and aligned
constexpr int LCHUNK = 16384;     // split path: matrices per chunk (bounds the workspace)
constexpr int SPLIT_MIN_L = 2048; // split path used from this batch size on (n <= 64; split_min_l)
// solver tuning flags (_x entry points): force the fused or the split n <= 64
// path; FLAG_TINYLOG shrinks the split path's rotation log so that every
// matrix takes the overflow fallback (a host-forced test of that path)
constexpr int FLAG_FUSED = 0x100, FLAG_SPLIT = 0x200, FLAG_TINYLOG = 0x400;
// split path: run the QL recurrence with d, e in LDS (eig_ql_lane_kernel) even
// where the register lockstep kernel applies (a host-side switch for tests)
constexpr int FLAG_QL_LDS = 0x800;
// split path phase selection: a call with some of these bits runs only those
// phases (tridiagonalisation + Q, QL recurrence, rotation apply + fallback);
// the phases hand over through the workspace, so consecutive calls with the
// same workspace and stream compose the full solve (batches of at most LCHUNK)
constexpr int FLAG_PH_TRI = 0x1000, FLAG_PH_QL = 0x2000, FLAG_PH_APPLY = 0x4000;
// FP32 64 < n <= 128: the register split path (two-wave tridiagonalisation,
// lane-per-matrix QL, 64-row-block apply) instead of the large-order one
constexpr int FLAG_REG128 = 0x8000;
constexpr int FLAG_PH_ALL = FLAG_PH_TRI | FLAG_PH_QL | FLAG_PH_APPLY;
// eigenpairs in the QL / Jacobi order instead of ascending (the reference's
// DO_SORT 0, eigen_GPU_batch.cu:5,10)
constexpr int FLAG_UNSORTED = 0x40000000;
// FP64 interface, mixed precision: FP32 solve of the rounded matrices, then
// FP64 refinement on MFMA (eig_refine.hpp) with the plain FP64 solver as the
// fallback for matrices the refinement does not accept
constexpr int SOLVER_MIXED = 3;
// with SOLVER_MIXED, 64 < n <= 256: the FP32 stage takes the large-order split
// path with the back-transformation as blocked reflectors on the matrix cores
// (FP16-split products, eig_wy.hpp) instead of forming Q on the vector units
constexpr int FLAG_MIXED_WY16 = 0x8000000;
// bits 28-29 of `solver`: refinement iterations (0: the default, 3)
constexpr int MIX_IT_SHIFT = 28;
constexpr int MIX_DEFAULT_ITERS = 3;
// refinement chunk (matrices whose FP64 work arrays are live at once)
constexpr int LCHUNK_REF = 2048;
// accepted when the last iteration's relative measures before the update are at
// most this (the update squares them).  From an FP32 start the first update
// leaves second-order terms of up to ~1e-5 (pairs just outside a cluster), the
// second ~1e-10, so with the default three iterations the bound is a check on
// the second update's result, and the third brings everything to FP64 level.
constexpr double MIX_ACCEPT = 1e-6;

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Split-path workspace regions for a chunk of lc matrices.
struct SplitLayout {
  size_t q, de, rot, sw, nsw, status, slot, total;
  SplitLayout(int lc, int n, int tb)
  {
    q = 0;
    de = q + al256((size_t)lc * n * n * tb);
    rot = de + al256((size_t)2 * n * lc * tb);
    sw = rot + al256((size_t)lc * rot_cap_of(n) * 2 * tb);
    nsw = sw + al256((size_t)lc * sw_cap_of(n) * 4);
    status = nsw + al256((size_t)lc * 4);
    slot = status + al256((size_t)lc * 4);
    total = slot + al256((size_t)lc * n * 4);
  }
};

// Large-order split path (orders beyond the register kernels): workspace
// regions for a chunk of lc matrices.
constexpr int LCHUNK_LARGE = 8192;
// bits 18-19 of `solver`: the split paths' chunk divided by 2, 4 or 8 (a smaller
// workspace for a longer call: the latency-bound QL recurrence runs once per chunk)
constexpr int CHUNK_SHIFT = 18;
int lchunk(int solver) { return LCHUNK >> ((solver >> CHUNK_SHIFT) & 3); }
int lchunk_large(int solver) { return LCHUNK_LARGE >> ((solver >> CHUNK_SHIFT) & 3); }
constexpr int SPLIT_LARGE_MIN_L = 512;
struct SplitLargeLayout {
  size_t wm, de, bw, rot, sw, nsw, status, slot, total;
  SplitLargeLayout(int lc, int n, int tb)
  {
    wm = 0;
    de = wm + al256((size_t)lc * n * n * tb);
    bw = de + al256((size_t)2 * n * lc * tb);
    rot = bw + al256((size_t)lc * n * tb);
    sw = rot + al256((size_t)lc * rot_cap_of(n) * 2 * tb);
    nsw = sw + al256((size_t)lc * sw_cap_of(n) * 4);
    status = nsw + al256((size_t)lc * 4);
    slot = status + al256((size_t)lc * 4);
    total = slot + al256((size_t)lc * n * 4);
  }
};

// Smallest batch for which the split path (tridiagonalisation, lane-per-matrix
// QL, replay) is the default for n <= 64 rather than the fused one-kernel
// path.  Measured on one MI355X, random symmetric matrices, L = 128..2048
// (split / fused ms): FP64 n = 16..32 the split path wins at every batch size
// (n = 32: 0.36 / 0.42 at L = 128, 0.38 / 0.48 at L = 1024; n = 20 within 2 %
// either way), FP32 from n = 28 (n = 32: 0.28 / 0.30 at L = 128; n = 20 and
// 24 the fused kernel by 3-16 %); n = 8 and 12 the fused kernel wins up to
// L = 2048 (FP64 n = 12: 0.110 / 0.105 at L = 2048); above n = 32 the fused
// kernel wins on small batches, up to L = 1024 for FP32 n >= 56 (n = 64:
// 0.92 / 0.97 at L = 1024) and to L = 2048 otherwise (FP64 n = 64: 1.35 / 1.20
// at L = 1024, 1.43 / 2.33 at L = 2048).
int split_min_l(int n, int tbytes)
{
  if (n <= 32 && n >= (tbytes == 8 ? 16 : 28)) return 1;
  if (tbytes == 4 && n >= 56) return 1024;
  return SPLIT_MIN_L;
}

bool use_split(int L, int n, int tbytes, int solver)
{
  const int base = solver & 0xff, flags = solver & ~0xff;
  if (n > 64 || (base != SOLVER_IMTQL2 && base != SOLVER_TQL2)) return false;
  if (flags & FLAG_FUSED) return false;
  if (flags & FLAG_SPLIT) return true;
  return L >= split_min_l(n, tbytes);
}

// Which kernel serves (n, element size): register-resident groups while a row
// of the matrix fits in the register file, the workgroup/global-memory kernel
// beyond.
bool use_wave_kernel(int n, int tbytes) { return tbytes == 4 ? n <= 128 : n <= 64; }

bool use_split_large(int L, int n, int tbytes, int solver)
{
  const int base = solver & 0xff, flags = solver & ~0xff;
  if (n <= 64 || n > GL_NMAX) return false;
  if (base != SOLVER_IMTQL2 && base != SOLVER_TQL2) return false;
  if (flags & FLAG_FUSED) return false;
  if (flags & FLAG_SPLIT) return true;
  // FP32 64 < n <= 128 also has the two-wave register kernel, faster for small
  // batches; measured at L = 8192: n = 96 25 ms split / 36 ms fused, n = 128 40 / 59 ms
  return L >= (use_wave_kernel(n, tbytes) ? SPLIT_MIN_L : SPLIT_LARGE_MIN_L);
}

// FP32 64 < n <= 128: the register split path is the default from
// SPLIT_MIN_L matrices on (measured at L = 8192: n = 128 12.3 ms, against
// 24.2 ms for the large-order split path and 59 ms for the fused kernel);
// FLAG_SPLIT still selects the large-order split path
bool use_split128(int L, int n, int tbytes, int solver)
{
  const int base = solver & 0xff, flags = solver & ~0xff;
  if (tbytes != 4 || n <= 64 || n > 128 || (base != SOLVER_IMTQL2 && base != SOLVER_TQL2)) return false;
  if (flags & (FLAG_FUSED | FLAG_SPLIT)) return false;
  if (flags & FLAG_REG128) return true;
  return L >= SPLIT_MIN_L;
}

int global_grid(int L) { return L < 1024 ? L : 1024; }

// Workspace: the register kernels keep everything in registers and LDS (a
// small fixed block keeps the reference's "wk == NULL means query" protocol
// yielding a non-NULL allocation); the global kernel needs one n x n
// accumulator per resident workgroup.
// Workspace of the path that `solver` (with its tuning bits) selects.  The
// size query of every entry point reports this for the bits the caller gave,
// never for a tuned choice: the reference ABI carries no workspace length, so
// a remembered choice is only taken when it fits what the query reported
// (tuned_solver).
size_t workspace_bytes(int L, int nm, int n, int m, int tbytes, int solver);

// FP64 refinement (eig_refine.hpp) of a chunk of lc matrices: six n x n FP64
// arrays (A, P = A X then E, X, X', S, G), the statistics step's per-matrix
// results, status, reasons and the largest first-order entry.
struct RefineLayout {
  size_t af, pe, x0, x1, sm, gm, inf, st, why, emx, total;
  int lc;
  RefineLayout(int L, int n, int solver = 0)
  {
    const int cap = LCHUNK_REF >> ((solver >> CHUNK_SHIFT) & 3);
    lc = L < cap ? L : cap;
    const size_t nn = (size_t)n * n;
    af = 0;
    pe = af + al256((size_t)lc * nn * 8);
    x0 = pe + al256((size_t)lc * nn * 8);
    x1 = x0 + al256((size_t)lc * nn * 8);
    sm = x1 + al256((size_t)lc * nn * 8);
    gm = sm + al256((size_t)lc * nn * 8);
    inf = gm + al256((size_t)lc * nn * 8);
    st = inf + al256((size_t)lc * rf_info_bytes(n));
    why = st + al256((size_t)lc * 4);
    emx = why + al256((size_t)lc * 4);
    total = emx + al256((size_t)lc * 8);
  }
};

// Mixed-precision path (FP64 interface): the FP32 copy of the batch, its
// eigenvalues and status, the FP32 solver's own workspace, then the refinement
// region.
// wy: the FP32 stage is the large-order split path with the blocked
// back-transformation, whose buffers follow the refinement region.
size_t wy_bytes(int lc, int n, int tb);
struct MixedLayout {
  size_t a32, w32, info32, ws32, ref, wyr, total;
  // cb: the caller's chunk bits, passed on to the FP32 stage and the refinement
  MixedLayout(int L, int n, bool wy, int cb = 0)
  {
    const size_t nn = (size_t)n * n;
    a32 = 0;
    w32 = a32 + al256((size_t)L * nn * 4);
    info32 = w32 + al256((size_t)L * n * 4);
    ws32 = info32 + al256((size_t)L * 4);
    ref = ws32 + al256(workspace_bytes(L, n, n, n, 4, (wy ? FLAG_SPLIT : 0) | cb));
    wyr = ref + al256(RefineLayout(L, n, cb).total);
    total = wyr + (wy ? wy_bytes(L < lchunk_large(cb) ? L : lchunk_large(cb), n, 4) : 0);
  }
};
bool mixed_wy(int n, int solver) { return (solver & FLAG_MIXED_WY16) && n > 64; }

size_t workspace_bytes(int L, int nm, int n, int m, int tbytes, int solver)
{
  (void)nm; (void)m;
  if ((solver & 0xff) == SOLVER_MIXED) {
    if (tbytes != 8 || n > 256) return 0;
    return WS_HEAD + MixedLayout(L, n, mixed_wy(n, solver), solver & (3 << CHUNK_SHIFT)).total;
  }
  if (use_split(L, n, tbytes, solver) || use_split128(L, n, tbytes, solver))
    return WS_HEAD + SplitLayout(L < lchunk(solver) ? L : lchunk(solver), n, tbytes).total;
  if (n > GL_NMAX) return 0;
  if (use_split_large(L, n, tbytes, solver))
    return WS_HEAD + SplitLargeLayout(L < lchunk_large(solver) ? L : lchunk_large(solver), n, tbytes).total;
  if (use_wave_kernel(n, tbytes)) return WS_HEAD;
  return WS_HEAD + (size_t)global_grid(L) * (size_t)n * (size_t)n * (size_t)tbytes;
}

template <typename T, int NB>
hipError_t launch_wave(int L, int nm, int n, T *a, T *w, int *info, int kop, hipStream_t s)
{
  constexpr int G = GroupShape<NB>::G, NW = GroupShape<NB>::NW;
  const long long blocks = ((long long)L + G - 1) / G;
  if (blocks > 0x7fffffffll) return hipErrorInvalidValue;
  hipLaunchKernelGGL((eig_wave_kernel<T, NB>), dim3((unsigned)blocks), dim3(64 * NW), 0, s,
                     L, nm, n, a, w, info, kop, (const int *)nullptr);
  return hipGetLastError();
}

template <typename T, int NB>
hipError_t launch_split(int L, int nm, int n, T *a, T *w, int *info, int solver, int kop, void *wk,
                        hipStream_t s)
{
  typedef typename vec2<T>::type T2;
  constexpr int G = GroupShape<NB>::G, NW = GroupShape<NB>::NW;
  const int lc = L < lchunk(solver) ? L : lchunk(solver);
  const SplitLayout lay(lc, n, (int)sizeof(T));
  char *base = (char *)wk + WS_HEAD;
  T *q = (T *)(base + lay.q), *de = (T *)(base + lay.de);
  T2 *rot = (T2 *)(base + lay.rot);
  int *sw = (int *)(base + lay.sw), *nsw = (int *)(base + lay.nsw);
  int *st = (int *)(base + lay.status), *slot = (int *)(base + lay.slot);
  const int ph = (solver & FLAG_PH_ALL) ? (solver & FLAG_PH_ALL) : FLAG_PH_ALL;
  if (ph != FLAG_PH_ALL && L > lc) return hipErrorInvalidValue;
  for (long long c0 = 0; c0 < L; c0 += lc) {
    const int cur = (int)((L - c0) < lc ? (L - c0) : lc);
    T *ac = a + (size_t)c0 * nm * n;
    T *wc = w + (size_t)c0 * n;
    int *ic = info ? info + c0 : nullptr;
    const unsigned gb = (unsigned)((cur + G - 1) / G), lb = (unsigned)((cur + 63) / 64);
    const int rcap = (solver & FLAG_TINYLOG) ? 1 : rot_cap_of(n), scap = sw_cap_of(n);
    if (ph & FLAG_PH_TRI)
      hipLaunchKernelGGL((eig_tri_kernel<T, NB>), dim3(gb), dim3(64 * NW), 0, s, cur, nm, n, ac, q, de, lc);
    if (ph & FLAG_PH_QL) {
      // registers hold d and e (2 NB values per lane): FP32 up to NB = 128, FP64 up to 64
      const bool reg = (solver & 0xff) == SOLVER_IMTQL2 && !(solver & FLAG_QL_LDS) && NB * sizeof(T) <= 512;
      if (reg)
        hipLaunchKernelGGL((eig_ql_reg_kernel<T, (NB * sizeof(T) <= 512 ? NB : 8)>), dim3(lb), dim3(64), 0, s,
                           cur, n, de, lc, rot, rcap, sw, scap, nsw, st, slot, wc, ic, kop);
      else
        hipLaunchKernelGGL((eig_ql_lane_kernel<T, NB>), dim3(lb), dim3(64), 0, s, cur, n, de, lc, rot, rcap,
                           sw, scap, nsw, st, slot, wc, ic, kop);
    }
    if (ph & FLAG_PH_APPLY) {
      // NB = 128: one workgroup per 64-row block (NW per matrix)
      hipLaunchKernelGGL((eig_apply_kernel<T, NB>), dim3(gb * NW), dim3(64), 0, s, cur, nm, n, ac, q, rot, rcap,
                         sw, scap, nsw, st, slot);
      hipLaunchKernelGGL((eig_wave_kernel<T, NB>), dim3(gb), dim3(64 * NW), 0, s, cur, nm, n, ac, wc, ic,
                         kop, (const int *)st);
    }
  }
  return hipGetLastError();
}

template <typename T, int NB>
hipError_t launch_jacobi(int L, int nm, int n, T *a, T *w, int *info, int kop, hipStream_t s)
{
  constexpr int G = 64 / NB;
  const long long blocks = ((long long)L + G - 1) / G;
  if (blocks > 0x7fffffffll) return hipErrorInvalidValue;
  hipLaunchKernelGGL((eig_jacobi_kernel<T, NB>), dim3((unsigned)blocks), dim3(64), 0, s, L, nm, n, a, w, info, kop);
  return hipGetLastError();
}

// Workgroup kernel tuning (bits of `solver`): workgroup size code at
// GL_NT_SHIFT (1: 256, 2: 512, 3: 1024 threads) and resident-workgroup count
// code at GL_GRID_SHIFT (g: 128 g workgroups, at most 1024); 0 picks the
// defaults.  Fewer, wider workgroups keep the matrices being worked on (and
// their accumulators) inside the on-chip caches.
constexpr int GL_NT_SHIFT = 16, GL_GRID_SHIFT = 20;
// large split path, tridiagonalisation shape (bits 24-25): 0: for n <= 256 the
// register-resident kernel (eig_tri_reg.hpp), beyond the default below; the
// workspace-resident kernels: 1: 256 threads, 2: 512 threads in 2 column
// groups, 3: 1024 threads in 4 column groups (2 for n > 256)
constexpr int TRI_CFG_SHIFT = 24;
// bit 26: for n <= 256, the full-storage tridiagonalisation instead of the
// default symmetric (lower-triangle) one
constexpr int TRI_FULL_BIT = 1 << 26;

template <typename T>
hipError_t launch_global(int L, int nm, int n, T *a, T *w, int *info, int solver, int kop, void *wk,
                         hipStream_t s)
{
  T *zws = (T *)((char *)wk + WS_HEAD);
  const int ntc = (solver >> GL_NT_SHIFT) & 3, gc = (solver >> GL_GRID_SHIFT) & 15;
  int grid = global_grid(L);
  if (gc) grid = (128 * gc < grid) ? 128 * gc : grid;
  const int base = kop;
  switch (ntc) {
  case 2:
    hipLaunchKernelGGL((eig_global_kernel<T, 512>), dim3(grid), dim3(512), 0, s, L, nm, n, a, w, info, base, zws,
                       (const int *)nullptr);
    break;
  case 3:
    hipLaunchKernelGGL((eig_global_kernel<T, 1024>), dim3(grid), dim3(1024), 0, s, L, nm, n, a, w, info, base,
                       zws, (const int *)nullptr);
    break;
  default:
    hipLaunchKernelGGL((eig_global_kernel<T, GL_THREADS>), dim3(grid), dim3(GL_THREADS), 0, s, L, nm, n, a, w,
                       info, base, zws, (const int *)nullptr);
  }
  return hipGetLastError();
}

// one workgroup per block of 64 RPL rows of each of the L matrices
template <typename T, int NCH, int CW, bool SKIPQ = false, int RPL = 1>
void launch_apply_large(int L, int nm, int n, T *a, const T *Wm, const T *bw,
                        const typename vec2<T>::type *rot, int rcap, const int *sw, int scap, const int *nsw,
                        const int *st, const int *slot, hipStream_t s)
{
  const unsigned grid = (unsigned)L * (unsigned)((n + 64 * RPL - 1) / (64 * RPL));
  hipLaunchKernelGGL((eig_apply_large_kernel<T, NCH, CW, false, SKIPQ, RPL>), dim3(grid), dim3(64 * NCH), 0, s, L,
                     nm, n, a, Wm, bw, rot, rcap, sw, scap, nsw, st, slot, UvSplitArgs{});
}

template <typename T>
size_t wy_bytes(int lc, int n, int tb);
hipError_t gemm_launch(int kind, const GemmArgs &g, hipStream_t s);

// Blocked back-transformation buffers for a chunk of lc matrices: scaled betas,
// U^T U and T per block, and the two panel products.
struct WyLayout {
  size_t bwy, g, t, p1, p2, total;
  WyLayout(int lc, int n, int tb)
  {
    bwy = 0;
    g = bwy + al256((size_t)lc * n * tb);
    t = g + al256((size_t)lc * WY_NB * WY_NB * tb);
    p1 = t + al256((size_t)lc * WY_NB * WY_NB * tb);
    p2 = p1 + al256((size_t)lc * WY_NB * n * tb);
    total = p2 + al256((size_t)lc * WY_NB * n * tb);
  }
};
size_t wy_bytes(int lc, int n, int tb) { return WyLayout(lc, n, tb).total; }

// V (ld nm) <- H_{n-1} ... H_1 V with the reflectors of the large-order
// tridiagonalisation (Wm, bw); matrices with status set keep V (eig_wy.hpp).
hipError_t wy_backtransform(int cur, int nm, int n, float *Wm, const float *bw, const int *st, float *V, char *wyb,
                            hipStream_t s)
{
  const WyLayout ly(cur, n, 4);
  float *bwy = (float *)(wyb + ly.bwy), *G = (float *)(wyb + ly.g), *Tm = (float *)(wyb + ly.t);
  float *P1 = (float *)(wyb + ly.p1), *P2 = (float *)(wyb + ly.p2);
  const long long cols = (long long)cur * n;
  hipLaunchKernelGGL(wy_prepare_kernel<float>, dim3((unsigned)((cols + 3) / 4 < 65536 ? (cols + 3) / 4 : 65536)),
                     dim3(256), 0, s, cur, n, Wm, bw, bwy, st);
  const long long nn = (long long)n * n;
  hipError_t err;
  for (int c0 = 0; c0 < n; c0 += WY_NB) {
    const int nb = (n - c0) < WY_NB ? (n - c0) : WY_NB;
    GemmArgs g{};
    g.batch = cur;
    g.alpha = 1.0;
    // G = U^T U
    g.M = g.N = nb; g.K = n; g.transa = 1; g.transb = 0;
    g.A = Wm + (size_t)c0 * n; g.lda = n; g.sa = nn;
    g.B = Wm + (size_t)c0 * n; g.ldb = n; g.sb = nn;
    g.C = G; g.ldc = WY_NB; g.sc = WY_NB * WY_NB;
    if ((err = gemm_launch(GK_F32, g, s)) != hipSuccess) return err;
    hipLaunchKernelGGL(wy_t_kernel<float>, dim3((unsigned)(cur < 65536 ? cur : 65536)), dim3(64), 0, s, cur, n, c0,
                       nb, (const float *)G, (const float *)bwy, Tm);
    // P1 = U^T V
    g.M = nb; g.N = n; g.K = n; g.transa = 1;
    g.B = V; g.ldb = nm; g.sb = (long long)nm * n;
    g.C = P1; g.ldc = WY_NB; g.sc = (long long)WY_NB * n;
    if ((err = gemm_launch(GK_F16X3, g, s)) != hipSuccess) return err;
    // P2 = T P1
    g.M = nb; g.N = n; g.K = nb; g.transa = 0;
    g.A = Tm; g.lda = WY_NB; g.sa = WY_NB * WY_NB;
    g.B = P1; g.ldb = WY_NB; g.sb = (long long)WY_NB * n;
    g.C = P2; g.ldc = WY_NB; g.sc = (long long)WY_NB * n;
    if ((err = gemm_launch(GK_F32, g, s)) != hipSuccess) return err;
    // V += U P2
    g.M = n; g.N = n; g.K = nb;
    g.A = Wm + (size_t)c0 * n; g.lda = n; g.sa = nn;
    g.B = P2; g.ldb = WY_NB; g.sb = (long long)WY_NB * n;
    g.D = V; g.ldd = nm; g.sd = (long long)nm * n; g.beta = 1.0;
    g.C = V; g.ldc = nm; g.sc = (long long)nm * n;
    if ((err = gemm_launch(GK_F16X3, g, s)) != hipSuccess) return err;
  }
  return hipSuccess;
}

// wyb (FP32 only, may be null): run the back-transformation as blocked
// reflectors on the matrix cores (the apply kernel then only replays the QL
// rotations from the identity) in a region of WyLayout(chunk, n, 4) bytes.
template <typename T>
hipError_t launch_split_large(int L, int nm, int n, T *a, T *w, int *info, int solver, int kop, void *wk,
                              hipStream_t s, char *wyb = nullptr)
{
  typedef typename vec2<T>::type T2;
  const int lc = L < lchunk_large(solver) ? L : lchunk_large(solver);
  const SplitLargeLayout lay(lc, n, (int)sizeof(T));
  char *base = (char *)wk + WS_HEAD;
  T *Wm = (T *)(base + lay.wm), *de = (T *)(base + lay.de), *bw = (T *)(base + lay.bw);
  T2 *rot = (T2 *)(base + lay.rot);
  int *sw = (int *)(base + lay.sw), *nsw = (int *)(base + lay.nsw);
  int *st = (int *)(base + lay.status), *slot = (int *)(base + lay.slot);
  const int rcap = (solver & FLAG_TINYLOG) ? 1 : rot_cap_of(n), scap = sw_cap_of(n);
  const int sv = kop;
  const int ph = (solver & FLAG_PH_ALL) ? (solver & FLAG_PH_ALL) : FLAG_PH_ALL;
  if (ph != FLAG_PH_ALL && L > lc) return hipErrorInvalidValue;
  for (long long c0 = 0; c0 < L; c0 += lc) {
    const int cur = (int)((L - c0) < lc ? (L - c0) : lc);
    T *ac = a + (size_t)c0 * nm * n;
    T *wc = w + (size_t)c0 * n;
    int *ic = info ? info + c0 : nullptr;
    const unsigned tg = (unsigned)(cur < 1024 ? cur : 1024);
    if (ph & FLAG_PH_TRI) {
      // tridiagonalisation shape: workgroup size / column groups and grid
      const int tc = (solver >> TRI_CFG_SHIFT) & 3, gc = (solver >> GL_GRID_SHIFT) & 15;
      const bool sym = n <= 256 && !(solver & TRI_FULL_BIT);
      // measured best at n = 128 and 256: symmetric form with 256 threads; full form with 512
      const int cfg = tc ? tc : (sym ? 1 : (n <= 256 ? 2 : 1));
      unsigned grid = tg;
      if (gc) grid = (unsigned)(128 * gc) < grid ? (unsigned)(128 * gc) : grid;
      if (sym && tc == 0) {
        // register-resident: 12 waves own the <= 36 tiles of n <= 256 (one
        // workgroup per CU), 4 waves the <= 10 tiles of n <= 128
        if (n <= 128)
          hipLaunchKernelGGL((eig_tri_reg_kernel<T, 4, 3, 3>), dim3(grid), dim3(256), 0, s, cur, nm, n, ac, Wm, de,
                             lc, bw);
        else
          hipLaunchKernelGGL((eig_tri_reg_kernel<T, 12, 3, 3>), dim3(grid), dim3(768), 0, s, cur, nm, n, ac, Wm, de,
                             lc, bw);
      } else if (sym) {
        if (cfg == 3)
          hipLaunchKernelGGL((eig_tri_large_kernel<T, 1024, TriNbp<T>::value, 256, 4, true>), dim3(grid), dim3(1024), 0, s, cur,
                             nm, n, ac, Wm, de, lc, bw);
        else if (cfg == 2)
          hipLaunchKernelGGL((eig_tri_large_kernel<T, 512, TriNbp<T>::value, 256, 2, true>), dim3(grid), dim3(512), 0, s, cur, nm,
                             n, ac, Wm, de, lc, bw);
        else
          hipLaunchKernelGGL((eig_tri_large_kernel<T, 256, TriNbp<T>::value, 256, 1, true>), dim3(grid), dim3(256), 0, s, cur, nm,
                             n, ac, Wm, de, lc, bw);
      } else if (n <= 256) {
        if (cfg == 3)
          hipLaunchKernelGGL((eig_tri_large_kernel<T, 1024, 8, 256, 4>), dim3(grid), dim3(1024), 0, s, cur, nm, n,
                             ac, Wm, de, lc, bw);
        else if (cfg == 2)
          hipLaunchKernelGGL((eig_tri_large_kernel<T, 512, 8, 256, 2>), dim3(grid), dim3(512), 0, s, cur, nm, n, ac,
                             Wm, de, lc, bw);
        else
          hipLaunchKernelGGL((eig_tri_large_kernel<T, 256, 8, 256, 1>), dim3(grid), dim3(256), 0, s, cur, nm, n, ac,
                             Wm, de, lc, bw);
      } else {
        if (cfg == 3)
          hipLaunchKernelGGL((eig_tri_large_kernel<T, 1024, 8, 512, 2>), dim3(grid), dim3(1024), 0, s, cur, nm, n,
                             ac, Wm, de, lc, bw);
        else
          hipLaunchKernelGGL((eig_tri_large_kernel<T, 256, 8, 512, 1>), dim3(grid), dim3(256), 0, s, cur, nm, n, ac,
                             Wm, de, lc, bw);
      }
    }
    if (!(ph & FLAG_PH_QL)) {
    } else if (n <= 256) {
      constexpr int NLANE = sizeof(T) == 8 ? 32 : 64;
      hipLaunchKernelGGL((eig_ql_lane_kernel<T, 256, NLANE>), dim3((cur + NLANE - 1) / NLANE), dim3(64), 0, s,
                         cur, n, de, lc, rot, rcap, sw, scap, nsw, st, slot, wc, ic, sv);
    } else {
      constexpr int NLANE = 8;
      hipLaunchKernelGGL((eig_ql_lane_kernel<T, 512, NLANE>), dim3((cur + NLANE - 1) / NLANE), dim3(64), 0, s,
                         cur, n, de, lc, rot, rcap, sw, scap, nsw, st, slot, wc, ic, sv);
    }
    if (!(ph & FLAG_PH_APPLY)) continue;
    constexpr int CWS = 32;  // chunk width
    if constexpr (sizeof(T) == 4) {
      if (wyb) {
        if (n <= 128)
          launch_apply_large<T, 128 / CWS, CWS, true>(cur, nm, n, ac, Wm, bw, rot, rcap, sw, scap, nsw, st, slot, s);
        else
          launch_apply_large<T, 256 / CWS, CWS, true>(cur, nm, n, ac, Wm, bw, rot, rcap, sw, scap, nsw, st, slot, s);
        hipError_t e2 = wy_backtransform(cur, nm, n, Wm, bw, st, ac, wyb, s);
        if (e2 != hipSuccess) return e2;
        hipLaunchKernelGGL((eig_global_kernel<T, GL_THREADS>), dim3(tg), dim3(GL_THREADS), 0, s, cur, nm, n, ac, wc,
                           ic, sv, Wm, (const int *)st);
        continue;
      }
    }
    // FP32 up to n = 256: two rows per lane (two chains per wave sharing each
    // rotation's LDS read)
    constexpr int RPL = sizeof(T) == 4 ? 2 : 1;
    if (n <= 128)
      launch_apply_large<T, 128 / CWS, CWS, false, RPL>(cur, nm, n, ac, Wm, bw, rot, rcap, sw, scap, nsw, st, slot, s);
    else if (n <= 256)
      launch_apply_large<T, 256 / CWS, CWS, false, RPL>(cur, nm, n, ac, Wm, bw, rot, rcap, sw, scap, nsw, st, slot, s);
    else if (n <= 384)
      launch_apply_large<T, 12, 32>(cur, nm, n, ac, Wm, bw, rot, rcap, sw, scap, nsw, st, slot, s);
    else
      launch_apply_large<T, 16, 32>(cur, nm, n, ac, Wm, bw, rot, rcap, sw, scap, nsw, st, slot, s);
    // overflow fallback: the workgroup solver on the flagged matrices, with
    // the (now free) tridiagonalisation slots as its accumulators
    hipLaunchKernelGGL((eig_global_kernel<T, GL_THREADS>), dim3(tg), dim3(GL_THREADS), 0, s, cur, nm, n, ac, wc,
                       ic, sv, Wm, (const int *)st);
  }
  return hipGetLastError();
}

// ---- batched products on the matrix cores ---------------------------------------

bool overlap(const void *a, size_t na, const void *b, size_t nb);

size_t gemm_extent(long long rows, long long cols, long long ld, long long stride, int batch)
{
  return (size_t)((batch - 1) * stride + (cols - 1) * ld + rows);
}

// Validates and launches eig_gemm_kernel<kind>: one workgroup per 64 x 64 tile.
hipError_t gemm_launch(int kind, const GemmArgs &g, hipStream_t s)
{
  if (kind < GK_F64 || kind > GK_F16X3) return hipErrorInvalidValue;
  if (g.M < 1 || g.N < 1 || g.K < 1 || g.batch < 0 || !g.A || !g.B || !g.C) return hipErrorInvalidValue;
  if (g.batch == 0) return hipSuccess;
  const long long ar = g.transa ? g.K : g.M, ac = g.transa ? g.M : g.K;
  const long long br = g.transb ? g.N : g.K, bc = g.transb ? g.K : g.N;
  if (g.lda < ar || g.ldb < br || g.ldc < g.M || (g.D && g.ldd < g.M)) return hipErrorInvalidValue;
  if (g.sa < 0 || g.sb < 0 || g.sc < 0 || g.sd < 0) return hipErrorInvalidValue;
  const int eb = kind == GK_F64 ? 8 : 4;
  const size_t ea = gemm_extent(ar, ac, g.lda, g.sa, g.batch) * eb;
  const size_t ebb = gemm_extent(br, bc, g.ldb, g.sb, g.batch) * eb;
  const size_t ec = gemm_extent(g.M, g.N, g.ldc, g.sc, g.batch) * eb;
  if (overlap(g.A, ea, g.C, ec) || overlap(g.B, ebb, g.C, ec)) return hipErrorInvalidValue;
  if (g.D && g.D != g.C && overlap(g.D, gemm_extent(g.M, g.N, g.ldd, g.sd, g.batch) * eb, g.C, ec))
    return hipErrorInvalidValue;
  if (g.D == g.C && g.ldd != g.ldc) return hipErrorInvalidValue;
  const long long tm = (g.M + GM_TILE - 1) / GM_TILE, tn = (g.N + GM_TILE - 1) / GM_TILE;
  long long tiles = tm * tn;
  if (g.sym) {
    if (g.sym != tm || tn % g.sym) return hipErrorInvalidValue;
    tiles = (tn / g.sym) * ((long long)g.sym * (g.sym + 1) / 2);
  }
  const long long total = tiles * g.batch;
  if (total > 0x7fffffffll) return hipErrorInvalidValue;
  switch (kind) {
  case GK_F64: hipLaunchKernelGGL((eig_gemm_kernel<GK_F64>), dim3((unsigned)total), dim3(GM_THREADS), 0, s, g); break;
  case GK_F32: hipLaunchKernelGGL((eig_gemm_kernel<GK_F32>), dim3((unsigned)total), dim3(GM_THREADS), 0, s, g); break;
  case GK_F16: hipLaunchKernelGGL((eig_gemm_kernel<GK_F16>), dim3((unsigned)total), dim3(GM_THREADS), 0, s, g); break;
  case GK_BF16: hipLaunchKernelGGL((eig_gemm_kernel<GK_BF16>), dim3((unsigned)total), dim3(GM_THREADS), 0, s, g); break;
  default: hipLaunchKernelGGL((eig_gemm_kernel<GK_F16X3>), dim3((unsigned)total), dim3(GM_THREADS), 0, s, g); break;
  }
  return hipGetLastError();
}

// square n x n products of a batch of packed (ld n) FP64 matrices
hipError_t gemm_sq64(int batch, int n, int ta, const double *A, const double *B, double beta, const double *D,
                     double *C, int sym, hipStream_t s)
{
  const long long nn = (long long)n * n;
  GemmArgs g{};
  g.M = g.N = g.K = n;
  g.batch = batch;
  g.transa = ta;
  g.transb = 0;
  g.A = A; g.lda = n; g.sa = nn;
  g.B = B; g.ldb = n; g.sb = nn;
  g.D = D; g.ldd = n; g.sd = D ? nn : 0;
  g.C = C; g.ldc = n; g.sc = nn;
  g.alpha = 1.0;
  g.beta = beta;
  g.sym = sym ? (n + GM_TILE - 1) / GM_TILE : 0;
  return gemm_launch(GK_F64, g, s);
}

template <typename T>
hipError_t run(int L, int nm, int n, int m, T *a, T *w, void *wk, hipStream_t s, int solver, int *info);

// The refinement iterations on one chunk (RefineLayout at rb): the full
// matrices are in af, X in x0, status / reasons initialised.  Returns the
// array holding the refined X.
hipError_t refine_chunk(int cur, int n, char *rb, const RefineLayout &ly, int iters, hipStream_t s, double **xout)
{
  double *af = (double *)(rb + ly.af), *pe = (double *)(rb + ly.pe);
  double *x = (double *)(rb + ly.x0), *xn = (double *)(rb + ly.x1);
  double *sm = (double *)(rb + ly.sm), *gm = (double *)(rb + ly.gm);
  char *inf = rb + ly.inf;
  int *st = (int *)(rb + ly.st), *why = (int *)(rb + ly.why);
  unsigned long long *emx = (unsigned long long *)(rb + ly.emx);
  const int nt = (n + 63) / 64;
  const long long fill = (long long)cur * nt * nt;
  hipError_t err;
  for (int it = 0; it < iters; ++it) {
    if ((err = gemm_sq64(cur, n, 0, af, x, 0.0, nullptr, pe, 0, s)) != hipSuccess) return err;   // P = A X
    if ((err = gemm_sq64(cur, n, 1, x, pe, 0.0, nullptr, sm, 1, s)) != hipSuccess) return err;   // S = X^T P
    if ((err = gemm_sq64(cur, n, 1, x, x, 0.0, nullptr, gm, 1, s)) != hipSuccess) return err;    // G = X^T X
    hipLaunchKernelGGL(eig_oa_stats_kernel, dim3((unsigned)(cur < 8192 ? cur : 8192)), dim3(RF_THREADS), 0, s, cur,
                       n, (const double *)sm, (const double *)gm, inf, ly.lc, st, why, emx);
    hipLaunchKernelGGL(eig_oa_fill_kernel, dim3((unsigned)(fill < (1 << 20) ? fill : (1 << 20))), dim3(256), 0, s,
                       cur, n, (const double *)sm, (const double *)gm, inf, ly.lc, pe, emx);
    if ((err = gemm_sq64(cur, n, 0, x, pe, 1.0, x, xn, 0, s)) != hipSuccess) return err;        // X' = X + X E
    double *tmp = x; x = xn; xn = tmp;
  }
  *xout = x;
  return hipGetLastError();
}

// Mixed precision for the FP64 interface: A rounded to FP32 (full symmetric,
// ld n), the FP32 solver on it (its default path and tuned choice), then per
// chunk the FP64 refinement (eig_refine.hpp), the store, and the plain FP64
// workgroup solver on the matrices the refinement did not accept.
hipError_t run_mixed(int L, int nm, int n, int m, double *a, double *w, void *wk, hipStream_t s, int solver,
                     int *info)
{
  if (n > 256) return hipErrorNotSupported;
  const int iters = ((solver >> MIX_IT_SHIFT) & 3) ? ((solver >> MIX_IT_SHIFT) & 3) : MIX_DEFAULT_ITERS;
  const int sorted = (solver & FLAG_UNSORTED) ? 0 : 1;
  const bool wy = mixed_wy(n, solver);
  const int cb = solver & (3 << CHUNK_SHIFT);
  const MixedLayout lay(L, n, wy, cb);
  const RefineLayout ly(L, n, cb);
  char *base = (char *)wk + WS_HEAD;
  char *rb = base + lay.ref;
  float *a32 = (float *)(base + lay.a32), *w32 = (float *)(base + lay.w32);
  int *info32 = (int *)(base + lay.info32);
  double *af = (double *)(rb + ly.af), *x0 = (double *)(rb + ly.x0), *sm = (double *)(rb + ly.sm);
  const double *lnew = (const double *)(rb + ly.inf) + (size_t)ly.lc * n;   // refined eigenvalues (RfInfo::lnew)
  int *st = (int *)(rb + ly.st), *why = (int *)(rb + ly.why);
  const unsigned long long *emx = (const unsigned long long *)(rb + ly.emx);
  const size_t nn = (size_t)n * n;
  const int nt = (n + 63) / 64, tiles = nt * (nt + 1) / 2;
  auto grid_of = [](long long work, long long cap) { return (unsigned)(work < cap ? work : cap); };
  hipLaunchKernelGGL((eig_symfull_kernel<double, float>), dim3(grid_of((long long)L * tiles, 65536)), dim3(256), 0,
                     s, L, n, (const double *)a, (long long)nm, (long long)nm * n, a32, (long long)n, (long long)nn);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  if (wy)
    err = launch_split_large<float>(L, n, n, a32, w32, info32, SOLVER_IMTQL2 | FLAG_SPLIT | cb,
                                    kopt_make(SOLVER_IMTQL2, false, 0), base + lay.ws32, s, base + lay.wyr);
  else
    err = run<float>(L, n, n, n, a32, w32, base + lay.ws32, s, SOLVER_IMTQL2 | cb, info32);
  if (err != hipSuccess) return err;
  for (long long c0 = 0; c0 < L; c0 += ly.lc) {
    const int cur = (int)((L - c0) < ly.lc ? (L - c0) : ly.lc);
    double *ac = a + (size_t)c0 * nm * n, *wc = w + (size_t)c0 * n;
    int *ic = info ? info + c0 : nullptr;
    hipLaunchKernelGGL((eig_symfull_kernel<double, double>), dim3(grid_of((long long)cur * tiles, 65536)), dim3(256),
                       0, s, cur, n, (const double *)ac, (long long)nm, (long long)nm * n, af, (long long)n,
                       (long long)nn);
    hipLaunchKernelGGL((eig_convert_kernel<float, double>), dim3(grid_of(((long long)cur * nn + 255) / 256, 16384)),
                       dim3(256), 0, s, (long long)cur * (long long)nn, (const float *)(a32 + (size_t)c0 * nn), x0);
    hipLaunchKernelGGL(eig_refine_init_kernel, dim3((cur + 255) / 256), dim3(256), 0, s, cur,
                       (const int *)(info32 + c0), st, why);
    double *x = nullptr;
    if ((err = refine_chunk(cur, n, rb, ly, iters, s, &x)) != hipSuccess) return err;
    hipLaunchKernelGGL(eig_refine_store_kernel, dim3(grid_of(cur, 4096)), dim3(RF_THREADS), 0, s, cur, nm, n, m,
                       (const double *)x, lnew, st, why, emx, MIX_ACCEPT, ac, wc, ic, sorted, 0);
    // the plain FP64 solver on the matrices left for it (S holds its accumulators)
    hipLaunchKernelGGL((eig_global_kernel<double, GL_THREADS>), dim3(grid_of(cur, 256)), dim3(GL_THREADS), 0, s, cur,
                       nm, n, ac, wc, ic, kopt_make(SOLVER_IMTQL2, !sorted, n - m), sm, (const int *)st);
    if ((err = hipGetLastError()) != hipSuccess) return err;
  }
  return hipSuccess;
}

// Refinement of given approximate eigenvectors (FP64): a holds the matrices
// (upper triangle, ld nm), x the approximate eigenvectors (ld ldx), refined in
// place with the eigenvalues into w, ascending; info[k] = 0 when accepted,
// else the reasons (RF_WHY_*) and x, w of that matrix hold the last iterate.
hipError_t run_refine(int L, int nm, int n, const double *a, double *x, int ldx, double *w, void *wk, hipStream_t s,
                      int iters, int *info)
{
  if (L < 1 || n < 1 || nm < n || ldx < n || !a || !x || !w || !wk) return hipErrorInvalidValue;
  if (n > 256) return hipErrorNotSupported;
  if (iters < 1 || iters > 8) return hipErrorInvalidValue;
  const RefineLayout ly(L, n);
  char *rb = (char *)wk + WS_HEAD;
  double *af = (double *)(rb + ly.af), *x0 = (double *)(rb + ly.x0);
  const double *lnew = (const double *)(rb + ly.inf) + (size_t)ly.lc * n;
  int *st = (int *)(rb + ly.st), *why = (int *)(rb + ly.why);
  const unsigned long long *emx = (const unsigned long long *)(rb + ly.emx);
  const size_t nn = (size_t)n * n;
  const int nt = (n + 63) / 64, tiles = nt * (nt + 1) / 2;
  auto grid_of = [](long long work, long long cap) { return (unsigned)(work < cap ? work : cap); };
  hipError_t err = hipSuccess;
  for (long long c0 = 0; c0 < L; c0 += ly.lc) {
    const int cur = (int)((L - c0) < ly.lc ? (L - c0) : ly.lc);
    const double *ac = a + (size_t)c0 * nm * n;
    double *xc = x + (size_t)c0 * ldx * n, *wc = w + (size_t)c0 * n;
    int *ic = info ? info + c0 : nullptr;
    hipLaunchKernelGGL((eig_symfull_kernel<double, double>), dim3(grid_of((long long)cur * tiles, 65536)), dim3(256),
                       0, s, cur, n, ac, (long long)nm, (long long)nm * n, af, (long long)n, (long long)nn);
    // X (ld ldx) -> x0 (ld n): the same tile copy with a full square source
    hipLaunchKernelGGL((eig_copy_cols_kernel<double>), dim3(grid_of((long long)cur * n, 65536)), dim3(256), 0, s,
                       cur, n, (const double *)xc, (long long)ldx, x0);
    hipLaunchKernelGGL(eig_refine_init_kernel, dim3((cur + 255) / 256), dim3(256), 0, s, cur, (const int *)nullptr,
                       st, why);
    double *xr = nullptr;
    if ((err = refine_chunk(cur, n, rb, ly, iters, s, &xr)) != hipSuccess) return err;
    hipLaunchKernelGGL(eig_refine_store_kernel, dim3(grid_of(cur, 4096)), dim3(RF_THREADS), 0, s, cur, ldx, n, n,
                       (const double *)xr, lnew, st, why, emx, MIX_ACCEPT, xc, wc, ic, 1, 1);
    if (ic)
      hipLaunchKernelGGL(eig_refine_info_kernel, dim3((cur + 255) / 256), dim3(256), 0, s, cur, (const int *)st,
                         (const int *)why, ic);
    if ((err = hipGetLastError()) != hipSuccess) return err;
  }
  return hipSuccess;
}

// ---- autotuned launch shapes ------------------------------------------------
// Tuning bits found by eigen_GPU_batch_tune_* per (device, element size, n,
// solver, power-of-two bucket of L); used when a call brings no tuning bits.

struct TuneKey {
  int dev, tb, n, solver, lb;
  bool operator<(const TuneKey &o) const
  {
    if (dev != o.dev) return dev < o.dev;
    if (tb != o.tb) return tb < o.tb;
    if (n != o.n) return n < o.n;
    if (solver != o.solver) return solver < o.solver;
    return lb < o.lb;
  }
};
std::mutex tune_mu;
std::map<TuneKey, int> tune_table;

int l_bucket(int L)
{
  int b = 0;
  while ((1 << (b + 1)) <= L && b < 30) ++b;
  return b;
}

TuneKey tune_key(int L, int n, int tb, int solver)
{
  int dev = 0;
  (void)hipGetDevice(&dev);
  return TuneKey{dev, tb, n, solver & 0xff, l_bucket(L)};
}

// The tuning bits a call runs with: the caller's own bits, else a remembered
// choice for this shape whose workspace is no larger than the one the size
// query reports for the call (workspace_bytes of the untuned solver), else
// none.  So a caller that sized wk before (or without) tuning never sees a
// path that needs more.
int tuned_solver(int L, int n, int tb, int solver)
{
  if (solver & ~0xff & ~FLAG_UNSORTED) return solver;  // the caller chose the path
  int bits = 0;
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    if (tune_table.empty()) return solver;
    auto it = tune_table.find(tune_key(L, n, tb, solver));
    if (it == tune_table.end()) return solver;
    bits = it->second;
  }
  if (workspace_bytes(L, n, n, n, tb, solver | bits) > workspace_bytes(L, n, n, n, tb, solver)) return solver;
  return solver | bits;
}

template <typename T>
hipError_t run(int L, int nm, int n, int m, T *a, T *w, void *wk, hipStream_t s, int solver, int *info)
{
  if (bad_args(L, nm, n, m)) return hipErrorInvalidValue;
  solver = tuned_solver(L, n, (int)sizeof(T), solver);
  const int base = solver & 0xff;
  if (base == SOLVER_MIXED) {
    if (wk == nullptr) return hipErrorInvalidValue;
    if constexpr (sizeof(T) == 8) return run_mixed(L, nm, n, m, a, w, wk, s, solver, info);
    return hipErrorNotSupported;
  }
  if (base != SOLVER_IMTQL2 && base != SOLVER_TQL2 && base != SOLVER_JACOBI) return hipErrorNotSupported;
  if (wk == nullptr) return hipErrorInvalidValue;
  // kernel options: solver, order of the output, eigenvector columns not stored
  const int kop = kopt_make(base, (solver & FLAG_UNSORTED) != 0, n - m);
  if ((solver & FLAG_PH_ALL) && !use_split(L, n, (int)sizeof(T), solver) && !use_split_large(L, n, (int)sizeof(T), solver) &&
      !use_split128(L, n, (int)sizeof(T), solver))
    return hipErrorInvalidValue;
  if constexpr (sizeof(T) == 4) {
    if (use_split128(L, n, (int)sizeof(T), solver)) return launch_split<T, 128>(L, nm, n, a, w, info, solver, kop, wk, s);
  }
  if (use_split(L, n, (int)sizeof(T), solver)) {
    if (n <= 8) return launch_split<T, 8>(L, nm, n, a, w, info, solver, kop, wk, s);
    if (n <= 16) return launch_split<T, 16>(L, nm, n, a, w, info, solver, kop, wk, s);
    if (n <= 32) return launch_split<T, 32>(L, nm, n, a, w, info, solver, kop, wk, s);
    return launch_split<T, 64>(L, nm, n, a, w, info, solver, kop, wk, s);
  }
  if (base == SOLVER_JACOBI) {
    if (n <= 8) return launch_jacobi<T, 8>(L, nm, n, a, w, info, kop, s);
    if (n <= 16) return launch_jacobi<T, 16>(L, nm, n, a, w, info, kop, s);
    if (n <= 32) return launch_jacobi<T, 32>(L, nm, n, a, w, info, kop, s);
    if (n <= 64) return launch_jacobi<T, 64>(L, nm, n, a, w, info, kop, s);
    return hipErrorNotSupported;
  }
  if (n <= 8) return launch_wave<T, 8>(L, nm, n, a, w, info, kop, s);
  if (n <= 16) return launch_wave<T, 16>(L, nm, n, a, w, info, kop, s);
  if (n <= 32) return launch_wave<T, 32>(L, nm, n, a, w, info, kop, s);
  if (n <= 64) return launch_wave<T, 64>(L, nm, n, a, w, info, kop, s);
  if (n > GL_NMAX) return hipErrorNotSupported;
  if (use_split_large(L, n, (int)sizeof(T), solver)) return launch_split_large<T>(L, nm, n, a, w, info, solver, kop, wk, s);
  if (use_wave_kernel(n, (int)sizeof(T)))
    return launch_wave<T, (sizeof(T) == 4 ? 128 : 64)>(L, nm, n, a, w, info, kop, s);
  return launch_global<T>(L, nm, n, a, w, info, solver, kop, wk, s);
}

// ---- autotuner --------------------------------------------------------------

// Candidate tuning bits for (n, element size, solver): the n <= 64 fused and
// split paths; for the large orders the workgroup kernel in two widths and the
// split path in three tridiagonalisation shapes and two grid sizes.
std::vector<int> tune_candidates(int L, int n, int tb, int solver)
{
  const int base = solver & 0xff;
  std::vector<int> c;
  if (base == SOLVER_JACOBI) return c;
  if (use_wave_kernel(n, tb)) {
    c.push_back(FLAG_FUSED);
    if (n <= 64) {
      c.push_back(FLAG_SPLIT);
      c.push_back(FLAG_SPLIT | FLAG_QL_LDS);
      c.push_back(FLAG_SPLIT | (1 << CHUNK_SHIFT));
      return c;
    }
    c.push_back(FLAG_REG128);
    c.push_back(FLAG_SPLIT | (1 << TRI_CFG_SHIFT));
    c.push_back(FLAG_SPLIT | (2 << TRI_CFG_SHIFT));
    c.push_back(FLAG_SPLIT | TRI_FULL_BIT | (2 << TRI_CFG_SHIFT));
    return c;
  }
  if (n > GL_NMAX) return c;
  c.push_back(FLAG_FUSED | (1 << GL_NT_SHIFT));
  c.push_back(FLAG_FUSED | (2 << GL_NT_SHIFT));
  // tridiagonalisation shapes: symmetric form 256 / 512 threads and the full
  // form with 512 threads (n <= 256); 256 / 1024 threads beyond
  const int cfs[2] = {1, n <= 256 ? 2 : 3};
  if (n <= 256) c.push_back(FLAG_SPLIT);
  for (int cf : cfs) {
    c.push_back(FLAG_SPLIT | (cf << TRI_CFG_SHIFT));
    if (L > 512) c.push_back(FLAG_SPLIT | (cf << TRI_CFG_SHIFT) | (4 << GL_GRID_SHIFT));
  }
  if (n <= 256) c.push_back(FLAG_SPLIT | TRI_FULL_BIT | (2 << TRI_CFG_SHIFT));
  return c;
}

size_t tune_bufsize(int L, int nm, int n, int m, int tb, int solver)
{
  size_t mx = WS_HEAD;
  for (int c : tune_candidates(L, n, tb, solver)) {
    const size_t b = workspace_bytes(L, nm, n, m, tb, (solver & 0xff) | c);
    mx = b > mx ? b : mx;
  }
  return mx;
}

template <typename T>
hipError_t run_tune(int L, int nm, int n, const T *a_src, T *a, T *w, void *wk, hipStream_t s, int solver, int *best,
                    float *best_ms, int max_report, int *cand, float *cand_ms, int *ncand)
{
  if (bad_args(L, nm, n, n) || !a_src || !a || !w || !wk || !best) return hipErrorInvalidValue;
  const std::vector<int> cands = tune_candidates(L, n, (int)sizeof(T), solver);
  if (cands.empty()) return hipErrorNotSupported;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return hipErrorOutOfMemory;
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return hipErrorOutOfMemory; }
  const size_t bytes = (size_t)L * nm * n * sizeof(T);
  // the workspace the default size query reports for this shape: only a
  // candidate that fits in it may be remembered for later default calls
  const size_t ws_default = workspace_bytes(L, nm, n, n, (int)sizeof(T), solver & 0xff);
  hipError_t err = hipSuccess;
  int bi = -1, bfit = -1;
  float bms = 0.0f, bfit_ms = 0.0f;
  int nrep = 0;
  for (size_t i = 0; i < cands.size() && err == hipSuccess; ++i) {
    const int sv = (solver & 0xff) | cands[i];
    float tmin = -1.0f;
    for (int rep = 0; rep < 3 && err == hipSuccess; ++rep) {   // one warm-up, two timed
      err = hipMemcpyAsync(a, a_src, bytes, hipMemcpyDeviceToDevice, s);
      if (err == hipSuccess) err = hipEventRecord(e0, s);
      if (err == hipSuccess) err = run<T>(L, nm, n, n, a, w, wk, s, sv, nullptr);
      if (err == hipSuccess) err = hipEventRecord(e1, s);
      if (err == hipSuccess) err = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
      if (err == hipSuccess && rep > 0) tmin = (tmin < 0.0f || ms < tmin) ? ms : tmin;
    }
    if (err != hipSuccess) break;
    if (cand && cand_ms && nrep < max_report) { cand[nrep] = cands[i]; cand_ms[nrep] = tmin; ++nrep; }
    if (bi < 0 || tmin < bms) { bi = (int)i; bms = tmin; }
    if (workspace_bytes(L, nm, n, n, (int)sizeof(T), sv) <= ws_default && (bfit < 0 || tmin < bfit_ms)) {
      bfit = (int)i;
      bfit_ms = tmin;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ncand) *ncand = nrep;
  if (err != hipSuccess) return err;
  *best = cands[bi];
  if (best_ms) *best_ms = bms;
  if (bfit >= 0) {
    std::lock_guard<std::mutex> lk(tune_mu);
    tune_table[tune_key(L, n, (int)sizeof(T), solver)] = cands[bfit];
  }
  // leave a holding the eigenvectors of the chosen configuration
  err = hipMemcpyAsync(a, a_src, bytes, hipMemcpyDeviceToDevice, s);
  if (err == hipSuccess) err = run<T>(L, nm, n, n, a, w, wk, s, (solver & 0xff) | cands[bi], nullptr);
  return err;
}

// ---- uneven batches ---------------------------------------------------------

// Size classes per element type: register groups of 8..64 lanes (and two-wave
// groups of 128 rows for FP32), the workgroup kernel beyond.
UvClasses uv_classes(int tbytes)
{
  UvClasses c{};
  const int f32[] = {8, 16, 32, 64, 128, GL_NMAX};
  const int f64[] = {8, 16, 32, 64, 256, GL_NMAX};
  const int *h = tbytes == 4 ? f32 : f64;
  c.ncls = 6;
  for (int i = 0; i < c.ncls; ++i) {
    c.hi[i] = h[i];
    c.g[i] = h[i] < 64 ? 64 / h[i] : 1;
  }
  return c;
}

int uv_global_grid(int L) { return L < 1024 ? L : 1024; }

// Split route of the uneven path: the class of orders (64, 128] (FP32) or
// (64, 256] (FP64) runs through the large split kernels (workgroup
// tridiagonalisation, lane-per-matrix QL, row-block apply) when it holds at
// least UV_SPLIT_MIN matrices, since the QL recurrence only pays off on many
// matrices at once; smaller classes keep the fused kernels.  The route works
// in chunks of cap matrices (workspace slots of order nmx).
constexpr int UV_SPLIT_MIN_F32 = 4096, UV_SPLIT_MIN_F64 = 1024;
int uv_split_nmx(int nmax, int tb)
{
  const int hi = tb == 4 ? 128 : 256;
  return nmax > 64 ? (nmax < hi ? nmax : hi) : 0;     // 0: no split route
}
// FP32: 65536 slots (17 GB at order 128), so that a class of up to 65536
// matrices pays the latency-bound QL recurrence once (two chunks of 32768 took
// 6.1 + 2.9 ms of it for 53000 matrices of orders 65..128)
int uv_split_cap(int L, int tb)
{
  const int ch = tb == 4 ? 65536 : 8192;
  return L < ch ? L : ch;
}
int uv_fallback_grid(int cap) { return cap < 256 ? cap : 256; }

struct UvSplitLayout {
  size_t wm, de, bw, rot, sw, nsw, status, slot, total;
  UvSplitLayout(int cap, int nmx, int tb)
  {
    wm = 0;
    de = wm + al256((size_t)cap * nmx * nmx * tb);
    bw = de + al256((size_t)2 * nmx * cap * tb);
    rot = bw + al256((size_t)cap * nmx * tb);
    sw = rot + al256((size_t)cap * rot_cap_of(nmx) * 2 * tb);
    nsw = sw + al256((size_t)cap * sw_cap_of(nmx) * sizeof(int));
    status = nsw + al256((size_t)cap * sizeof(int));
    slot = status + al256((size_t)cap * sizeof(int));
    total = slot + al256((size_t)cap * nmx * sizeof(int));
  }
};

struct UvLayout {
  size_t ctl, list, aoff, woff, z, split, total;
  UvLayout(int L, int nmax, int tb)
  {
    ctl = 0;
    list = ctl + al256(sizeof(int) * 4 * UV_MAXCLS);
    aoff = list + al256((size_t)uv_list_len(L) * sizeof(int));
    woff = aoff + al256((size_t)L * sizeof(long long));
    z = woff + al256((size_t)L * sizeof(long long));
    const bool glob = !use_wave_kernel(nmax, tb);
    const int nmx = uv_split_nmx(nmax, tb), cap = uv_split_cap(L, tb);
    const size_t zg = glob ? (size_t)uv_global_grid(L) * nmax * nmax * tb : 0;
    const size_t zs = nmx ? (size_t)uv_fallback_grid(cap) * nmx * nmx * tb : 0;
    split = z + al256(zg > zs ? zg : zs);
    total = split + (nmx ? UvSplitLayout(cap, nmx, tb).total : 0);
  }
};

// Workgroups that fit on the device at once for a kernel (cached per device
// and kernel; the first call per pair queries the runtime).
int resident_blocks(const void *kernel, int threads)
{
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find({dev, kernel});
  if (it != cache.end()) return it->second;
  int cus = 0, per = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1) per = 1;
  const int nb = cus * per;
  cache[{dev, kernel}] = nb;
  return nb;
}

template <typename T, int NB>
void launch_wave_list(int *ctl, const int *list, const int *nv, const long long *ao, const long long *wo, T *a,
                      T *w, int *info, int solver, hipStream_t s)
{
  constexpr int NW = GroupShape<NB>::NW;
  const void *kern = (const void *)eig_wave_list_kernel<T, NB>;
  const int blocks = resident_blocks(kern, 64 * NW);
  hipLaunchKernelGGL((eig_wave_list_kernel<T, NB>), dim3(blocks), dim3(64 * NW), 0, s, ctl, list, nv, ao, wo,
                     a, w, info, solver);
}

// The split route's kernels for one class, chunk by chunk (chunks beyond the
// class's entries find nothing to do; see uv_split_entry).
template <typename T>
void run_uneven_split(int L, int nmx, int *cc, const int *list, const int *nv, const long long *ao,
                      const long long *wo, T *a, T *w, int *info, int solver, char *base, T *zws, hipStream_t s)
{
  typedef typename vec2<T>::type T2;
  const int tb = (int)sizeof(T);
  const int cap = uv_split_cap(L, tb);
  const UvSplitLayout lay(cap, nmx, tb);
  T *Wm = (T *)(base + lay.wm), *de = (T *)(base + lay.de), *bw = (T *)(base + lay.bw);
  T2 *rot = (T2 *)(base + lay.rot);
  int *sw = (int *)(base + lay.sw), *nsw = (int *)(base + lay.nsw);
  int *st = (int *)(base + lay.status), *slot = (int *)(base + lay.slot);
  const int rcap = (solver & FLAG_TINYLOG) ? 1 : rot_cap_of(nmx), scap = sw_cap_of(nmx);
  const int sv = solver & 0xff;
  const int min_count = (solver & FLAG_SPLIT) ? 1 : (tb == 4 ? UV_SPLIT_MIN_F32 : UV_SPLIT_MIN_F64);
  constexpr int QMAX = sizeof(T) == 4 ? 128 : 256, NLANE = sizeof(T) == 4 ? 64 : 32;
  for (long long c0 = 0; c0 < L; c0 += cap) {
    const UvSplitArgs u{cc, list, nv, ao, wo, (int)c0, cap, min_count, nmx};
    if constexpr (sizeof(T) == 4) {
      // FP32 orders 65..128: the register split kernels (two waves per
      // matrix in the tridiagonalisation, one workgroup per 64-row block in
      // the replay), Q in the Wm slots
      hipLaunchKernelGGL((eig_tri_kernel<T, 128, true>), dim3(cap), dim3(128), 0, s, cap, 0, 0, (const T *)a, Wm,
                         de, cap, u);
      if (sv == SOLVER_IMTQL2 && !(solver & FLAG_QL_LDS))
        hipLaunchKernelGGL((eig_ql_reg_kernel<T, 128, true>), dim3((cap + 63) / 64), dim3(64), 0, s, cap, 0,
                           (const T *)de, cap, rot, rcap, sw, scap, nsw, st, slot, w, info, sv, u);
      else
        hipLaunchKernelGGL((eig_ql_lane_kernel<T, QMAX, NLANE, true>), dim3((cap + NLANE - 1) / NLANE), dim3(64), 0,
                           s, cap, 0, (const T *)de, cap, rot, rcap, sw, scap, nsw, st, slot, w, info, sv, u);
      hipLaunchKernelGGL((eig_apply_kernel<T, 128, true>), dim3((unsigned)cap * 2), dim3(64), 0, s, cap, 0, 0, a,
                         (const T *)Wm, (const T2 *)rot, rcap, (const int *)sw, scap, (const int *)nsw,
                         (const int *)st, (const int *)slot, u);
    } else {
      // FP64 orders 65..256: the large-order split kernels
      const int nrb = (nmx + 63) >> 6;
      hipLaunchKernelGGL((eig_tri_large_list_kernel<T, 256, TriNbp<T>::value, 256, 1, true>),
                         dim3(cap < 1024 ? cap : 1024), dim3(256), 0, s, u, (const T *)a, Wm, de, bw);
      hipLaunchKernelGGL((eig_ql_lane_kernel<T, QMAX, NLANE, true>), dim3((cap + NLANE - 1) / NLANE), dim3(64), 0,
                         s, cap, 0, (const T *)de, cap, rot, rcap, sw, scap, nsw, st, slot, w, info, sv, u);
      if (nmx <= 128)
        hipLaunchKernelGGL((eig_apply_large_kernel<T, 4, 32, true>), dim3((unsigned)cap * nrb), dim3(256), 0, s, cap,
                           0, 0, a, (const T *)Wm, (const T *)bw, (const T2 *)rot, rcap, (const int *)sw, scap,
                           (const int *)nsw, (const int *)st, (const int *)slot, u);
      else
        hipLaunchKernelGGL((eig_apply_large_kernel<T, 8, 32, true>), dim3((unsigned)cap * nrb), dim3(512), 0, s, cap,
                           0, 0, a, (const T *)Wm, (const T *)bw, (const T2 *)rot, rcap, (const int *)sw, scap,
                           (const int *)nsw, (const int *)st, (const int *)slot, u);
    }
    hipLaunchKernelGGL((eig_global_fallback_list_kernel<T>), dim3(uv_fallback_grid(cap)), dim3(GL_THREADS), 0, s,
                       u, (const int *)st, a, w, info, sv, zws);
  }
}

// The class (32, 64] on the n <= 64 split kernels in list form (one wave per
// matrix in the tridiagonalisation and the apply, the LDS QL kernel with an
// order per lane), in the slots of the split route's region laid out for
// order 64 (they fit in the region of any larger order)
template <typename T>
void run_uneven_split64(int L, int *cc, const int *list, const int *nv, const long long *ao, const long long *wo,
                        T *a, T *w, int *info, int solver, char *base, T *zws, hipStream_t s)
{
  typedef typename vec2<T>::type T2;
  const int tb = (int)sizeof(T), nmx = 64;
  const int cap = uv_split_cap(L, tb);
  const UvSplitLayout lay(cap, nmx, tb);
  T *Wm = (T *)(base + lay.wm), *de = (T *)(base + lay.de);
  T2 *rot = (T2 *)(base + lay.rot);
  int *sw = (int *)(base + lay.sw), *nsw = (int *)(base + lay.nsw);
  int *st = (int *)(base + lay.status), *slot = (int *)(base + lay.slot);
  const int rcap = (solver & FLAG_TINYLOG) ? 1 : rot_cap_of(nmx), scap = sw_cap_of(nmx);
  const int sv = solver & 0xff;
  const int min_count = (solver & FLAG_SPLIT) ? 1 : (tb == 4 ? UV_SPLIT_MIN_F32 : UV_SPLIT_MIN_F64);
  constexpr int NLANE = sizeof(T) == 4 ? 64 : 32;
  for (long long c0 = 0; c0 < L; c0 += cap) {
    const UvSplitArgs u{cc, list, nv, ao, wo, (int)c0, cap, min_count, nmx};
    hipLaunchKernelGGL((eig_tri_kernel<T, 64, true>), dim3(cap), dim3(64), 0, s, cap, 0, 0, (const T *)a, Wm, de,
                       cap, u);
    if (sv == SOLVER_IMTQL2 && !(solver & FLAG_QL_LDS))
      hipLaunchKernelGGL((eig_ql_reg_kernel<T, 64, true>), dim3((cap + 63) / 64), dim3(64), 0, s, cap, 0,
                         (const T *)de, cap, rot, rcap, sw, scap, nsw, st, slot, w, info, sv, u);
    else
      hipLaunchKernelGGL((eig_ql_lane_kernel<T, 64, NLANE, true>), dim3((cap + NLANE - 1) / NLANE), dim3(64), 0, s,
                         cap, 0, (const T *)de, cap, rot, rcap, sw, scap, nsw, st, slot, w, info, sv, u);
    hipLaunchKernelGGL((eig_apply_kernel<T, 64, true>), dim3(cap), dim3(64), 0, s, cap, 0, 0, a, (const T *)Wm,
                       (const T2 *)rot, rcap, (const int *)sw, scap, (const int *)nsw, (const int *)st,
                       (const int *)slot, u);
    hipLaunchKernelGGL((eig_global_fallback_list_kernel<T>), dim3(uv_fallback_grid(cap)), dim3(GL_THREADS), 0, s,
                       u, (const int *)st, a, w, info, sv, zws);
  }
}

template <typename T>
hipError_t run_uneven(int L, int nmax, const int *nv, const long long *aoff, const long long *woff, T *a, T *w,
                      void *wk, hipStream_t s, int solver, int *info)
{
  if (L < 1 || nmax < 1 || nmax > GL_NMAX || !nv || !a || !w || !wk) return hipErrorInvalidValue;
  if ((aoff == nullptr) != (woff == nullptr)) return hipErrorInvalidValue;
  const int base = solver & 0xff;
  if (base != SOLVER_IMTQL2 && base != SOLVER_TQL2) return hipErrorNotSupported;
  const int tb = (int)sizeof(T);
  const UvLayout lay(L, nmax, tb);
  char *ws = (char *)wk + WS_HEAD;
  int *ctl = (int *)(ws + lay.ctl), *list = (int *)(ws + lay.list);
  long long *aws = (long long *)(ws + lay.aoff), *wws = (long long *)(ws + lay.woff);
  T *zws = (T *)(ws + lay.z);
  const UvClasses cls = uv_classes(tb);
  hipLaunchKernelGGL(uv_bin_kernel, dim3(1), dim3(UV_THREADS), 0, s, L, nmax, nv, aws, wws,
                     aoff == nullptr ? 1 : 0, list, ctl, info, cls);
  const long long *ao = aoff ? aoff : aws, *wo = woff ? woff : wws;
  const int nmx = uv_split_nmx(nmax, tb);
  for (int c = 0; c < cls.ncls; ++c) {
    const int lo = c == 0 ? 0 : cls.hi[c - 1];
    if (lo >= nmax) break;
    int *cc = ctl + 4 * c;
    // split route: the class beyond order 64, and the class (32, 64] when the
    // workspace holds the route's region (nmax > 64)
    const bool split = nmx && (lo == 64 || lo == 32) && !(solver & FLAG_FUSED);
    const int min_count = (solver & FLAG_SPLIT) ? 1 : (tb == 4 ? UV_SPLIT_MIN_F32 : UV_SPLIT_MIN_F64);
    if (split) hipLaunchKernelGGL(uv_claim_kernel, dim3(1), dim3(64), 0, s, cc, min_count);
    switch (cls.hi[c]) {
    case 8: launch_wave_list<T, 8>(cc, list, nv, ao, wo, a, w, info, base, s); break;
    case 16: launch_wave_list<T, 16>(cc, list, nv, ao, wo, a, w, info, base, s); break;
    case 32: launch_wave_list<T, 32>(cc, list, nv, ao, wo, a, w, info, base, s); break;
    case 64: launch_wave_list<T, 64>(cc, list, nv, ao, wo, a, w, info, base, s); break;
    case 128: launch_wave_list<T, (sizeof(T) == 4 ? 128 : 64)>(cc, list, nv, ao, wo, a, w, info, base, s); break;
    default:
      hipLaunchKernelGGL((eig_global_list_kernel<T>), dim3(uv_global_grid(L)), dim3(GL_THREADS), 0, s, cc,
                         list, nv, ao, wo, a, w, info, base, zws, (long long)nmax * nmax);
    }
    if (split && lo == 64)
      run_uneven_split<T>(L, nmx, cc, list, nv, ao, wo, a, w, info, solver, ws + lay.split, zws, s);
    else if (split)
      run_uneven_split64<T>(L, cc, list, nv, ao, wo, a, w, info, solver, ws + lay.split, zws, s);
  }
  return hipGetLastError();
}

// ---- warm-started solve (n <= 64) ------------------------------------------------
// Workspace per chunk of lc matrices: the full matrix, later B = Q^T A Q and its
// Jacobi eigenvectors (one region), Q, A Q then V = Q Z_B (one region), the
// Jacobi status and the acceptance status.
constexpr int LCHUNK_WARM = 16384;
struct WarmLayout {
  size_t ab, q, v, ji, st, total;
  int lc;
  WarmLayout(int L, int n, int tb)
  {
    lc = L < LCHUNK_WARM ? L : LCHUNK_WARM;
    const size_t mat = (size_t)lc * n * n * tb;
    ab = 0;
    q = al256(mat);
    v = q + al256(mat);
    ji = v + al256(mat);
    st = ji + al256((size_t)lc * 4);
    total = st + al256((size_t)lc * 4);
  }
};

size_t warm_bytes(int L, int n, int tb)
{
  if (L < 1 || n < 1 || n > 64 || (tb != 4 && tb != 8)) return 0;
  return WS_HEAD + WarmLayout(L, n, tb).total;
}

template <typename T, int NB>
void launch_jacobi_capped(int L, int n, T *b, T *w, int *jinfo, int maxsw, int *nsweeps, hipStream_t s)
{
  constexpr int G = 64 / NB;
  hipLaunchKernelGGL((eig_jacobi_kernel<T, NB>), dim3((unsigned)((L + G - 1) / G)), dim3(64), 0, s, L, n, n, b, w,
                     jinfo, kopt_make(SOLVER_JACOBI, false, 0), maxsw, nsweeps);
}

template <typename T, int NB>
void launch_wave_only(int L, int nm, int n, T *a, T *w, int *info, int kop, const int *only, hipStream_t s)
{
  constexpr int G = GroupShape<NB>::G, NW = GroupShape<NB>::NW;
  hipLaunchKernelGGL((eig_wave_kernel<T, NB>), dim3((unsigned)((L + G - 1) / G)), dim3(64 * NW), 0, s, L, nm, n, a,
                     w, info, kop, only);
}

// q0: L approximate eigenbases in a's layout (ld nm).  Per chunk: full A, Q =
// qr(q0), A Q, B = Q^T (A Q) (matrix cores: FP32 operands for FP32, FP64 for
// FP64), Jacobi on B with at most maxsw sweeps, V = Q Z_B, then V stored for the
// accepted matrices and the plain solver run on the others.
template <typename T>
hipError_t run_warm(int L, int nm, int n, int m, T *a, const T *q0, T *w, void *wk, hipStream_t s, int maxsw,
                    int *info, int *sweeps)
{
  if (bad_args(L, nm, n, m) || !a || !q0 || !w || !wk) return hipErrorInvalidValue;
  if (n > 64) return hipErrorNotSupported;
  if (maxsw < 1 || maxsw > JAC_MAX_SWEEP) return hipErrorInvalidValue;
  const int tb = (int)sizeof(T);
  const WarmLayout lay(L, n, tb);
  char *base = (char *)wk + WS_HEAD;
  T *ab = (T *)(base + lay.ab), *q = (T *)(base + lay.q), *v = (T *)(base + lay.v);
  int *ji = (int *)(base + lay.ji), *st = (int *)(base + lay.st);
  const size_t nn = (size_t)n * n;
  const int kind = tb == 8 ? GK_F64 : GK_F32;
  const int kop = kopt_make(SOLVER_IMTQL2, false, n - m);
  hipError_t err = hipSuccess;
  for (long long c0 = 0; c0 < L; c0 += lay.lc) {
    const int cur = (int)((L - c0) < lay.lc ? (L - c0) : lay.lc);
    T *ac = a + (size_t)c0 * nm * n, *wc = w + (size_t)c0 * n;
    const T *qc = q0 + (size_t)c0 * nm * n;
    int *ic = info ? info + c0 : nullptr, *sc = sweeps ? sweeps + c0 : nullptr;
    hipLaunchKernelGGL((eig_symfull_kernel<T, T>), dim3((unsigned)(cur < 65536 ? cur : 65536)), dim3(256), 0, s, cur,
                       n, (const T *)ac, (long long)nm, (long long)nm * n, ab, (long long)n, (long long)nn);
    hipLaunchKernelGGL((eig_qr_kernel<T>), dim3((unsigned)(cur < 65536 ? cur : 65536)), dim3(64), 0, s, cur, nm, n,
                       qc, q, (long long)n, (long long)nn);
    GemmArgs g{};
    g.M = g.N = g.K = n;
    g.batch = cur;
    g.lda = g.ldb = g.ldc = n;
    g.sa = g.sb = g.sc = (long long)nn;
    g.alpha = 1.0;
    g.A = ab; g.B = q; g.C = v;                      // A Q
    if ((err = gemm_launch(kind, g, s)) != hipSuccess) return err;
    g.transa = 1; g.A = q; g.B = v; g.C = ab;        // B = Q^T (A Q)
    if ((err = gemm_launch(kind, g, s)) != hipSuccess) return err;
    if (n <= 8) launch_jacobi_capped<T, 8>(cur, n, ab, wc, ji, maxsw, sc, s);
    else if (n <= 16) launch_jacobi_capped<T, 16>(cur, n, ab, wc, ji, maxsw, sc, s);
    else if (n <= 32) launch_jacobi_capped<T, 32>(cur, n, ab, wc, ji, maxsw, sc, s);
    else launch_jacobi_capped<T, 64>(cur, n, ab, wc, ji, maxsw, sc, s);
    g.transa = 0; g.A = q; g.B = ab; g.C = v;        // V = Q Z_B
    if ((err = gemm_launch(kind, g, s)) != hipSuccess) return err;
    hipLaunchKernelGGL((eig_warm_store_kernel<T>), dim3((unsigned)(cur < 16384 ? cur : 16384)), dim3(256), 0, s, cur,
                       nm, n, m, (const T *)v, (const int *)ji, st, ac, ic);
    if (n <= 8) launch_wave_only<T, 8>(cur, nm, n, ac, wc, ic, kop, st, s);
    else if (n <= 16) launch_wave_only<T, 16>(cur, nm, n, ac, wc, ic, kop, st, s);
    else if (n <= 32) launch_wave_only<T, 32>(cur, nm, n, ac, wc, ic, kop, st, s);
    else launch_wave_only<T, 64>(cur, nm, n, ac, wc, ic, kop, st, s);
    if ((err = hipGetLastError()) != hipSuccess) return err;
  }
  return hipSuccess;
}

template <typename T>
hipError_t run_qr(int L, int nm, int n, const T *a, T *q, hipStream_t s)
{
  if (L < 0 || n < 1 || n > 64 || nm < n || !a || !q) return hipErrorInvalidValue;
  if (L == 0) return hipSuccess;
  hipLaunchKernelGGL((eig_qr_kernel<T>), dim3((unsigned)(L < 65536 ? L : 65536)), dim3(64), 0, s, L, nm, n, a, q,
                     (long long)nm, (long long)nm * n);
  return hipGetLastError();
}

// ---- warm-start network (eig_warmnet.hpp) ------------------------------------------

bool wn_ok(int n, int h) { return n >= 1 && n <= 64 && h >= 1 && h <= 4096; }

// activations of a batch of B: features, the two hidden layers, the output (also
// its gradient) and the two hidden gradients
struct WnLayout {
  size_t x, h1, h2, y, d2, d1, total;
  WnLayout(int B, const WnDims &d)
  {
    x = 0;
    h1 = x + al256((size_t)d.din * B * 4);
    h2 = h1 + al256((size_t)d.h * B * 4);
    y = h2 + al256((size_t)d.h * B * 4);
    d2 = y + al256((size_t)d.dout * B * 4);
    d1 = d2 + al256((size_t)d.h * B * 4);
    total = d1 + al256((size_t)d.h * B * 4);
  }
};

unsigned wn_grid(long long work) { return (unsigned)((work + 255) / 256 < 65536 ? (work + 255) / 256 : 65536); }

// one product C = op(A) op(B) on the matrix cores (FP32 operands)
hipError_t wn_gemm(int M, int N, int K, int ta, int tb, const float *A, int lda, const float *B, int ldb, float *C,
                   int ldc, hipStream_t s)
{
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K; g.batch = 1;
  g.transa = ta; g.transb = tb;
  g.A = A; g.lda = lda;
  g.B = B; g.ldb = ldb;
  g.C = C; g.ldc = ldc;
  g.alpha = 1.0;
  return gemm_launch(GK_F32, g, s);
}

hipError_t wn_forward(const WnDims &d, int B, const float *p, char *ws, const WnLayout &ly, hipStream_t s)
{
  float *x = (float *)(ws + ly.x), *h1 = (float *)(ws + ly.h1), *h2 = (float *)(ws + ly.h2);
  float *y = (float *)(ws + ly.y);
  hipError_t err;
  if ((err = wn_gemm(d.h, B, d.din, 1, 0, p + d.w1, d.din, x, d.din, h1, d.h, s)) != hipSuccess) return err;
  hipLaunchKernelGGL(wn_bias_act_kernel, dim3(wn_grid((long long)d.h * B)), dim3(256), 0, s, d.h, (long long)B, h1,
                     (long long)d.h, p + d.b1, 1);
  if ((err = wn_gemm(d.h, B, d.h, 1, 0, p + d.w2, d.h, h1, d.h, h2, d.h, s)) != hipSuccess) return err;
  hipLaunchKernelGGL(wn_bias_act_kernel, dim3(wn_grid((long long)d.h * B)), dim3(256), 0, s, d.h, (long long)B, h2,
                     (long long)d.h, p + d.b2, 1);
  if ((err = wn_gemm(d.dout, B, d.h, 1, 0, p + d.w3, d.h, h2, d.h, y, d.dout, s)) != hipSuccess) return err;
  hipLaunchKernelGGL(wn_bias_act_kernel, dim3(wn_grid((long long)d.dout * B)), dim3(256), 0, s, d.dout,
                     (long long)B, y, (long long)d.dout, p + d.b3, 0);
  return hipGetLastError();
}

hipError_t wn_run_forward(int L, int nm, int n, int h, const float *p, const float *a, float *q0, int batch, void *wk,
                          hipStream_t s)
{
  if (!wn_ok(n, h) || L < 0 || nm < n || batch < 1 || !p || !a || !q0 || !wk) return hipErrorInvalidValue;
  const WnDims d = wn_dims(n, h);
  const WnLayout ly(batch, d);
  char *ws = (char *)wk;
  for (long long c0 = 0; c0 < L; c0 += batch) {
    const int cur = (int)((L - c0) < batch ? (L - c0) : batch);
    hipLaunchKernelGGL(wn_features_kernel, dim3((unsigned)(cur < 65536 ? cur : 65536)), dim3(256), 0, s, cur, nm, n,
                       a + (size_t)c0 * nm * n, (float *)(ws + ly.x));
    hipError_t err = wn_forward(d, cur, p, ws, ly, s);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(wn_store_q_kernel, dim3(wn_grid((long long)cur * n * n)), dim3(256), 0, s, cur, nm, n,
                       (const float *)(ws + ly.y), q0 + (size_t)c0 * nm * n);
  }
  return hipGetLastError();
}

void wn_rowsum(int rows, long long cols, const float *z, long long ld, float *db, hipStream_t s)
{
  const int split = cols >= 4096 ? 64 : (cols >= 256 ? 8 : 1);
  hipLaunchKernelGGL(wn_rowsum_kernel, dim3((rows + 255) / 256, split), dim3(256), 0, s, rows, cols, z, ld, db, split);
}

hipError_t wn_run_grad(int B, int nm, int n, int h, const float *p, const float *a, const float *v, float *g,
                       float *loss, void *wk, hipStream_t s)
{
  if (!wn_ok(n, h) || B < 1 || nm < n || !p || !a || !v || !g || !loss || !wk) return hipErrorInvalidValue;
  const WnDims d = wn_dims(n, h);
  const WnLayout ly(B, d);
  char *ws = (char *)wk;
  float *x = (float *)(ws + ly.x), *h1 = (float *)(ws + ly.h1), *h2 = (float *)(ws + ly.h2);
  float *y = (float *)(ws + ly.y), *d2 = (float *)(ws + ly.d2), *d1 = (float *)(ws + ly.d1);
  hipLaunchKernelGGL(wn_features_kernel, dim3((unsigned)(B < 65536 ? B : 65536)), dim3(256), 0, s, B, nm, n, a, x);
  hipError_t err = wn_forward(d, B, p, ws, ly, s);
  if (err != hipSuccess) return err;
  // the bias gradients accumulate: clear them (and the loss) first
  hipLaunchKernelGGL(wn_fill_kernel, dim3(1), dim3(256), 0, s, 1ll, loss, 0.0f);
  hipLaunchKernelGGL(wn_fill_kernel, dim3(wn_grid(d.h)), dim3(256), 0, s, (long long)d.h, g + d.b1, 0.0f);
  hipLaunchKernelGGL(wn_fill_kernel, dim3(wn_grid(d.h)), dim3(256), 0, s, (long long)d.h, g + d.b2, 0.0f);
  hipLaunchKernelGGL(wn_fill_kernel, dim3(wn_grid(d.dout)), dim3(256), 0, s, (long long)d.dout, g + d.b3, 0.0f);
  hipLaunchKernelGGL(wn_mse_grad_kernel, dim3(wn_grid((long long)B * n)), dim3(256), 0, s, B, nm, n, y, v, loss);
  // layer 3
  if ((err = wn_gemm(d.h, d.dout, B, 0, 1, h2, d.h, y, d.dout, g + d.w3, d.h, s)) != hipSuccess) return err;
  wn_rowsum(d.dout, B, y, d.dout, g + d.b3, s);
  if ((err = wn_gemm(d.h, B, d.dout, 0, 0, p + d.w3, d.h, y, d.dout, d2, d.h, s)) != hipSuccess) return err;
  hipLaunchKernelGGL(wn_relu_grad_kernel, dim3(wn_grid((long long)d.h * B)), dim3(256), 0, s, (long long)d.h * B, d2,
                     (const float *)h2);
  // layer 2
  if ((err = wn_gemm(d.h, d.h, B, 0, 1, h1, d.h, d2, d.h, g + d.w2, d.h, s)) != hipSuccess) return err;
  wn_rowsum(d.h, B, d2, d.h, g + d.b2, s);
  if ((err = wn_gemm(d.h, B, d.h, 0, 0, p + d.w2, d.h, d2, d.h, d1, d.h, s)) != hipSuccess) return err;
  hipLaunchKernelGGL(wn_relu_grad_kernel, dim3(wn_grid((long long)d.h * B)), dim3(256), 0, s, (long long)d.h * B, d1,
                     (const float *)h1);
  // layer 1
  if ((err = wn_gemm(d.din, d.h, B, 0, 1, x, d.din, d1, d.h, g + d.w1, d.din, s)) != hipSuccess) return err;
  wn_rowsum(d.h, B, d1, d.h, g + d.b1, s);
  return hipGetLastError();
}

// ---- interleaved layout ---------------------------------------------------------

size_t interleaved_bytes(int L, int nm, int n, int m, int tb, int solver)
{
  if (bad_args(L, nm, n, m) || (tb != 4 && tb != 8)) return 0;
  const size_t inner = workspace_bytes(L, nm, n, m, tb, solver);
  if (inner == 0) return 0;
  return al256((size_t)L * nm * n * tb) + al256((size_t)L * n * tb) + inner;
}

hipError_t transpose_launch(long long R, long long C, const void *in, void *out, int tb, hipStream_t s)
{
  const long long tiles = ((R + 63) / 64) * ((C + 63) / 64);
  const unsigned grid = (unsigned)(tiles < 65536 ? tiles : 65536);
  if (tb == 8)
    hipLaunchKernelGGL((eig_transpose_kernel<double>), dim3(grid), dim3(256), 0, s, R, C, (const double *)in,
                       (double *)out);
  else
    hipLaunchKernelGGL((eig_transpose_kernel<float>), dim3(grid), dim3(256), 0, s, R, C, (const float *)in,
                       (float *)out);
  return hipGetLastError();
}

// Interleaved batch (element (r, c) of matrix k at a[k + L (r + c nm)],
// eigenvalue j at w[k + L j]): transposed into the batch-strided layout in the
// workspace, solved there, transposed back.
template <typename T>
hipError_t run_interleaved(int L, int nm, int n, int m, T *a, T *w, void *wk, hipStream_t s, int solver, int *info)
{
  if (bad_args(L, nm, n, m) || !a || !w || !wk) return hipErrorInvalidValue;
  const size_t inner = workspace_bytes(L, nm, n, m, (int)sizeof(T), solver);
  if (inner == 0) return hipErrorNotSupported;
  char *base = (char *)wk;
  T *as = (T *)base;
  T *ws = (T *)(base + al256((size_t)L * nm * n * sizeof(T)));
  void *wk2 = base + al256((size_t)L * nm * n * sizeof(T)) + al256((size_t)L * n * sizeof(T));
  const long long e = (long long)nm * n;
  hipError_t err = transpose_launch(e, L, a, as, (int)sizeof(T), s);
  if (err == hipSuccess) err = run<T>(L, nm, n, m, as, ws, wk2, s, solver, info);
  if (err == hipSuccess) err = transpose_launch(L, e, as, a, (int)sizeof(T), s);
  if (err == hipSuccess) err = transpose_launch(L, n, ws, w, (int)sizeof(T), s);
  return err;
}

// ---- accuracy check -----------------------------------------------------------

template <typename T>
hipError_t run_check(int L, int nm, int n, int m, const T *a, const T *w, const T *z, double *out, int into_a,
                     hipStream_t s)
{
  if (L < 1 || n < 1 || nm < n || m < 1 || m > n || n > GL_NMAX || !a || !w || !z) return hipErrorInvalidValue;
  // the report (four floats) written over each a_k must fit inside a_k, or it
  // would overwrite the next matrix (or run past the end of a)
  if (into_a && (size_t)nm * n * sizeof(T) < 4 * sizeof(float)) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)(L < 4096 ? L : 4096);
  hipLaunchKernelGGL((eig_check_kernel<T>), dim3(grid), dim3(CK_THREADS), 0, s, L, nm, n, m, a, w, z, out, into_a);
  return hipGetLastError();
}

// The reference's eigen_GPU_check_RUN (eigen_GPU_check.cu:143-182): report
// written over each a_k, then read back and the failing matrices printed.
template <typename T>
void check_entry(int L, int nm, int n, int m, T *a, T *w, T *z, hipStream_t s)
{
  if ((size_t)nm * n * sizeof(T) < 4 * sizeof(float)) {
    fprintf(stderr, "eigen_GPU_check: a matrix of %d x %d elements of %d bytes cannot hold the 16-byte report; "
            "use eigen_GPU_check_x_*\n", nm, n, (int)sizeof(T));
    return;
  }
  if (run_check<T>(L, nm, n, m, a, w, z, nullptr, 1, s) != hipSuccess) return;
  std::vector<float> rep((size_t)L * 4);
  const size_t pitch = (size_t)nm * n * sizeof(T);
  if (hipMemcpy2DAsync(rep.data(), 4 * sizeof(float), a, pitch, 4 * sizeof(float), L, hipMemcpyDeviceToHost, s) !=
      hipSuccess)
    return;
  if (hipStreamSynchronize(s) != hipSuccess) return;
  for (int id = 0; id < L; ++id) {
    const float *ans = &rep[(size_t)id * 4];
    if (ans[0] != 0.0f)
      printf("[%06d] Accuracy error in the relative resisdual (||Ax-dx||_F=%le/%le/%le)\n", id, (double)ans[0],
             (double)ans[2], (double)ans[3]);
    if (ans[1] != 0.0f)
      printf("[%06d] Orthonormality error (||ZZ-I||_F=%le/%le/%le)\n", id, (double)ans[1], (double)ans[2],
             (double)ans[3]);
  }
}

// ---- products around the solver (MFMA) --------------------------------------

// threads per workgroup: one wave per 32 x 32 output block of the triangle, at most 4
int sd_threads(int n)
{
  const int nb = (n + 31) >> 5, nblk = nb * (nb + 1) / 2;
  return 64 * (nblk < 4 ? nblk : 4);
}

bool overlap(const void *a, size_t na, const void *b, size_t nb)
{
  const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
  return x < y + nb && y < x + na;
}

template <typename T>
hipError_t run_symdmt(int L, int nm, int n, const T *z, const T *d, int fn, T p, T scale, T *c, hipStream_t s)
{
  if (L < 0 || n < 1 || n > GL_NMAX || nm < n || !z || !d || !c) return hipErrorInvalidValue;
  if (L == 0) return hipSuccess;
  const size_t bytes = (size_t)L * nm * n * sizeof(T);
  if (overlap(z, bytes, c, bytes)) return hipErrorInvalidValue;
  if (n <= 64) {
    // one wave per matrix, four waves per workgroup
    const unsigned grid = (unsigned)((L + 3) / 4 < 2048 ? (L + 3) / 4 : 2048);
    switch ((n + 15) >> 4) {
    case 1: hipLaunchKernelGGL((eig_symdmt_wave_kernel<T, 1>), dim3(grid), dim3(256), 0, s, L, nm, n, z, d, fn, p, scale, c); break;
    case 2: hipLaunchKernelGGL((eig_symdmt_wave_kernel<T, 2>), dim3(grid), dim3(256), 0, s, L, nm, n, z, d, fn, p, scale, c); break;
    case 3: hipLaunchKernelGGL((eig_symdmt_wave_kernel<T, 3>), dim3(grid), dim3(256), 0, s, L, nm, n, z, d, fn, p, scale, c); break;
    default: hipLaunchKernelGGL((eig_symdmt_wave_kernel<T, 4>), dim3(grid), dim3(256), 0, s, L, nm, n, z, d, fn, p, scale, c);
    }
    return hipGetLastError();
  }
  const unsigned grid = (unsigned)(L < 4096 ? L : 4096);
  hipLaunchKernelGGL((eig_symdmt_kernel<T>), dim3(grid), dim3(sd_threads(n)), 0, s, L, nm, n, z, d, fn, p, scale,
                     c);
  return hipGetLastError();
}

template <typename T>
hipError_t run_gram(int L, int nm, int n, int np, const T *y, int ldy, const T *rinv, int rstride, T alpha, T *a,
                    hipStream_t s)
{
  if (L < 0 || n < 1 || n > GL_NMAX || nm < n || np < 1 || ldy < np || rstride < 0 || !y || !a)
    return hipErrorInvalidValue;
  if (L == 0) return hipSuccess;
  if (overlap(y, (size_t)L * ldy * n * sizeof(T), a, (size_t)L * nm * n * sizeof(T))) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)(L < 4096 ? L : 4096);
  const bool vec = (ldy % 4 == 0) && (np % 4 == 0) && ((uintptr_t)y % 16 == 0) &&
                   (!rinv || ((uintptr_t)rinv % 16 == 0 && rstride % 4 == 0));
  if (n <= 64) {
    const unsigned gw = (unsigned)((L + 3) / 4 < 2048 ? (L + 3) / 4 : 2048);
    const int nt = (n + 15) >> 4;
#define GRAM_WAVE(NT, V)                                                                                          \
  hipLaunchKernelGGL((eig_gram_wave_kernel<T, NT, V>), dim3(gw), dim3(256), 0, s, L, nm, n, np, y, ldy, rinv,     \
                     rstride, alpha, a)
    if (vec) {
      if (nt == 1) GRAM_WAVE(1, true); else if (nt == 2) GRAM_WAVE(2, true);
      else if (nt == 3) GRAM_WAVE(3, true); else GRAM_WAVE(4, true);
    } else {
      if (nt == 1) GRAM_WAVE(1, false); else if (nt == 2) GRAM_WAVE(2, false);
      else if (nt == 3) GRAM_WAVE(3, false); else GRAM_WAVE(4, false);
    }
#undef GRAM_WAVE
    return hipGetLastError();
  }
  if (vec)
    hipLaunchKernelGGL((eig_gram_kernel<T, true>), dim3(grid), dim3(sd_threads(n)), 0, s, L, nm, n, np, y, ldy, rinv,
                       rstride, alpha, a);
  else
    hipLaunchKernelGGL((eig_gram_kernel<T, false>), dim3(grid), dim3(sd_threads(n)), 0, s, L, nm, n, np, y, ldy,
                       rinv, rstride, alpha, a);
  return hipGetLastError();
}

// Launch shape of the first kernel of the path run() takes, in the terms of
// the reference's eigen_GPU_batch_get_Launch_params (eigen_GPU_batch.hpp:23-102):
// p[0] lanes per matrix group, p[1] workgroups, p[2] matrices (groups) per
// workgroup, p[3] threads per workgroup, p[4] dynamic LDS bytes passed at launch
// (0: every kernel here sizes its LDS statically), p[5] L2 bytes of the device.
hipError_t launch_params(int L, int n, int tb, int solver, int *p)
{
  for (int i = 0; i < 6; ++i) p[i] = 0;
  if (L < 1 || n < 1 || n > GL_NMAX || (tb != 4 && tb != 8)) return hipErrorInvalidValue;
  solver = tuned_solver(L, n, tb, solver);
  const int base = solver & 0xff;
  auto nb_of = [](int n_) { return n_ <= 8 ? 8 : n_ <= 16 ? 16 : n_ <= 32 ? 32 : 64; };
  auto grp = [&](int nb, int nmat, int threads) {
    const int g = nb < 64 ? 64 / nb : 1;
    p[0] = nb;
    p[2] = g;
    p[3] = threads;
    p[1] = (nmat + g - 1) / g;
  };
  if (tb == 4 && use_split128(L, n, tb, solver)) {
    grp(128, L < lchunk(solver) ? L : lchunk(solver), 128);
  } else if (use_split(L, n, tb, solver)) {
    grp(nb_of(n), L < lchunk(solver) ? L : lchunk(solver), 64);
  } else if (base == SOLVER_JACOBI) {
    if (n > 64) return hipErrorNotSupported;
    grp(nb_of(n), L, 64);
  } else if (n <= 64) {
    grp(nb_of(n), L, 64);
  } else if (use_split_large(L, n, tb, solver)) {
    const int lc = L < lchunk_large(solver) ? L : lchunk_large(solver);
    const int tc = (solver >> TRI_CFG_SHIFT) & 3, gc = (solver >> GL_GRID_SHIFT) & 15;
    const bool sym = n <= 256 && !(solver & TRI_FULL_BIT);
    const int cfg = tc ? tc : (sym ? 1 : (n <= 256 ? 2 : 1));
    int grid = lc < 1024 ? lc : 1024;
    if (gc && 128 * gc < grid) grid = 128 * gc;
    const int threads = cfg == 3 ? 1024 : cfg == 2 ? 512 : 256;
    p[0] = threads; p[1] = grid; p[2] = 1; p[3] = threads;
  } else if (use_wave_kernel(n, tb)) {
    grp(128, L, 128);
  } else {
    const int ntc = (solver >> GL_NT_SHIFT) & 3, gc = (solver >> GL_GRID_SHIFT) & 15;
    int grid = global_grid(L);
    if (gc && 128 * gc < grid) grid = 128 * gc;
    const int threads = ntc == 2 ? 512 : ntc == 3 ? 1024 : GL_THREADS;
    p[0] = threads; p[1] = grid; p[2] = 1; p[3] = threads;
  }
  int dev = 0, l2 = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&l2, hipDeviceAttributeL2CacheSize, dev) == hipSuccess)
    p[5] = l2;
  return hipSuccess;
}

template <typename T>
void abi_entry(int L, int nm, int n, int m, T *a, T *w, T *wk, hipStream_t s)
{
  if (wk == nullptr) {
    // size query (eigen_GPU_batch.cu:331-339): always a full size_t
    if (w != nullptr) {
      const size_t lw = bad_args(L, nm, n, m) ? 0 : workspace_bytes(L, nm, n, m, (int)sizeof(T), 0);
      memcpy((void *)w, &lw, sizeof(size_t));
    }
    return;
  }
  (void)run<T>(L, nm, n, m, a, w, wk, s, SOLVER_IMTQL2, nullptr);
}

} // namespace

extern "C" {

void eigen_GPU_batch_DP(const int L, const int nm, const int n, const int m, double *a, double *w,
                        double *wk, const gpuStream_t stream)
{
  abi_entry<double>(L, nm, n, m, a, w, wk, stream);
}

void eigen_GPU_batch_FP(const int L, const int nm, const int n, const int m, float *a, float *w,
                        float *wk, const gpuStream_t stream)
{
  abi_entry<float>(L, nm, n, m, a, w, wk, stream);
}

size_t eigen_GPU_batch_bufsize(int L, int nm, int n, int m, int dtype_bytes, int solver)
{
  if (bad_args(L, nm, n, m) || (dtype_bytes != 4 && dtype_bytes != 8)) return 0;
  return workspace_bytes(L, nm, n, m, dtype_bytes, solver);
}

int eigen_GPU_batch_FP_x(int L, int nm, int n, int m, float *a, float *w, void *wk, hipStream_t stream,
                         int solver, int *info)
{
  return (int)run<float>(L, nm, n, m, a, w, wk, stream, solver, info);
}

int eigen_GPU_batch_DP_x(int L, int nm, int n, int m, double *a, double *w, void *wk, hipStream_t stream,
                         int solver, int *info)
{
  return (int)run<double>(L, nm, n, m, a, w, wk, stream, solver, info);
}

size_t eigen_GPU_batch_uneven_bufsize(int L, int nmax, int dtype_bytes, int solver)
{
  (void)solver;
  if (L < 1 || nmax < 1 || nmax > GL_NMAX || (dtype_bytes != 4 && dtype_bytes != 8)) return 0;
  return WS_HEAD + UvLayout(L, nmax, dtype_bytes).total;
}

int eigen_GPU_batch_uneven_FP(int L, int nmax, const int *n, const long long *aoff, const long long *woff,
                              float *a, float *w, void *wk, hipStream_t stream, int solver, int *info)
{
  return (int)run_uneven<float>(L, nmax, n, aoff, woff, a, w, wk, stream, solver, info);
}

int eigen_GPU_batch_uneven_DP(int L, int nmax, const int *n, const long long *aoff, const long long *woff,
                              double *a, double *w, void *wk, hipStream_t stream, int solver, int *info)
{
  return (int)run_uneven<double>(L, nmax, n, aoff, woff, a, w, wk, stream, solver, info);
}

void eigen_GPU_check_DP(const int L, const int nm, const int n, const int m, double *a_, double *w_, double *z_,
                        const gpuStream_t stream)
{
  check_entry<double>(L, nm, n, m, a_, w_, z_, stream);
}

void eigen_GPU_check_FP(const int L, const int nm, const int n, const int m, float *a_, float *w_, float *z_,
                        const gpuStream_t stream)
{
  check_entry<float>(L, nm, n, m, a_, w_, z_, stream);
}

int eigen_GPU_check_x_DP(int L, int nm, int n, int m, const double *a, const double *w, const double *z,
                         double *out, hipStream_t stream)
{
  if (!out) return (int)hipErrorInvalidValue;
  return (int)run_check<double>(L, nm, n, m, a, w, z, out, 0, stream);
}

int eigen_GPU_check_x_FP(int L, int nm, int n, int m, const float *a, const float *w, const float *z,
                         double *out, hipStream_t stream)
{
  if (!out) return (int)hipErrorInvalidValue;
  return (int)run_check<float>(L, nm, n, m, a, w, z, out, 0, stream);
}

size_t eigen_GPU_batch_tune_bufsize(int L, int nm, int n, int dtype_bytes, int solver)
{
  if (bad_args(L, nm, n, n) || (dtype_bytes != 4 && dtype_bytes != 8)) return 0;
  return tune_bufsize(L, nm, n, n, dtype_bytes, solver);
}

int eigen_GPU_batch_tune_FP(int L, int nm, int n, const float *a_src, float *a, float *w, void *wk,
                            hipStream_t stream, int solver, int *best, float *best_ms, int max_report, int *cand,
                            float *cand_ms, int *ncand)
{
  return (int)run_tune<float>(L, nm, n, a_src, a, w, wk, stream, solver, best, best_ms, max_report, cand, cand_ms,
                              ncand);
}

int eigen_GPU_batch_tune_DP(int L, int nm, int n, const double *a_src, double *a, double *w, void *wk,
                            hipStream_t stream, int solver, int *best, float *best_ms, int max_report, int *cand,
                            float *cand_ms, int *ncand)
{
  return (int)run_tune<double>(L, nm, n, a_src, a, w, wk, stream, solver, best, best_ms, max_report, cand, cand_ms,
                               ncand);
}

int eigen_GPU_batch_tune_set(int L, int n, int dtype_bytes, int solver, int bits)
{
  if (L < 1 || n < 1 || n > GL_NMAX || (dtype_bytes != 4 && dtype_bytes != 8) || (bits & 0xff) ||
      (bits & (FLAG_PH_ALL | FLAG_TINYLOG)))
    return (int)hipErrorInvalidValue;
  std::lock_guard<std::mutex> lk(tune_mu);
  const TuneKey key = tune_key(L, n, dtype_bytes, solver);
  if (bits == 0) tune_table.erase(key);
  else tune_table[key] = bits;
  return (int)hipSuccess;
}

void eigen_GPU_batch_tune_reset(void)
{
  std::lock_guard<std::mutex> lk(tune_mu);
  tune_table.clear();
}

int eigen_GPU_batch_symdmt_FP(int L, int nm, int n, const float *z, const float *d, float *c, hipStream_t stream)
{
  return (int)run_symdmt<float>(L, nm, n, z, d, 0, 0.0f, 1.0f, c, stream);
}

int eigen_GPU_batch_symdmt_DP(int L, int nm, int n, const double *z, const double *d, double *c,
                              hipStream_t stream)
{
  return (int)run_symdmt<double>(L, nm, n, z, d, 0, 0.0, 1.0, c, stream);
}

int eigen_GPU_batch_matpow_FP(int L, int nm, int n, const float *z, const float *w, float p, float scale, float *c,
                              hipStream_t stream)
{
  return (int)run_symdmt<float>(L, nm, n, z, w, 1, p, scale, c, stream);
}

int eigen_GPU_batch_matpow_DP(int L, int nm, int n, const double *z, const double *w, double p, double scale,
                              double *c, hipStream_t stream)
{
  return (int)run_symdmt<double>(L, nm, n, z, w, 1, p, scale, c, stream);
}

int eigen_GPU_batch_gram_FP(int L, int nm, int n, int p, const float *y, int ldy, const float *rinv, int rstride,
                            float alpha, float *a, hipStream_t stream)
{
  return (int)run_gram<float>(L, nm, n, p, y, ldy, rinv, rstride, alpha, a, stream);
}

int eigen_GPU_batch_gram_DP(int L, int nm, int n, int p, const double *y, int ldy, const double *rinv, int rstride,
                            double alpha, double *a, hipStream_t stream)
{
  return (int)run_gram<double>(L, nm, n, p, y, ldy, rinv, rstride, alpha, a, stream);
}

int eigen_GPU_batch_launch_params(int L, int n, int dtype_bytes, int solver, int *params)
{
  if (!params) return (int)hipErrorInvalidValue;
  return (int)launch_params(L, n, dtype_bytes, solver, params);
}

size_t eigen_GPU_warmnet_params(int n, int hidden)
{
  return wn_ok(n, hidden) ? (size_t)wn_dims(n, hidden).count : 0;
}

size_t eigen_GPU_warmnet_bufsize(int batch, int n, int hidden)
{
  if (!wn_ok(n, hidden) || batch < 1) return 0;
  return WnLayout(batch, wn_dims(n, hidden)).total;
}

int eigen_GPU_warmnet_init(int n, int hidden, float *params, unsigned long long seed, hipStream_t stream)
{
  if (!wn_ok(n, hidden) || !params) return hipErrorInvalidValue;
  const WnDims d = wn_dims(n, hidden);
  hipLaunchKernelGGL(wn_init_kernel, dim3(wn_grid(d.count)), dim3(256), 0, stream, d, params, seed);
  return hipGetLastError();
}

int eigen_GPU_warmnet_forward(int L, int nm, int n, int hidden, const float *params, const float *a, float *q0,
                              int batch, void *wk, hipStream_t stream)
{
  return wn_run_forward(L, nm, n, hidden, params, a, q0, batch, wk, stream);
}

int eigen_GPU_warmnet_grad(int B, int nm, int n, int hidden, const float *params, const float *a, const float *v,
                           float *grad, float *loss, void *wk, hipStream_t stream)
{
  return wn_run_grad(B, nm, n, hidden, params, a, v, grad, loss, wk, stream);
}

int eigen_GPU_warmnet_adam(long long count, float *params, const float *grad, float *m, float *v, int step, float lr,
                           hipStream_t stream)
{
  if (count < 1 || step < 1 || !params || !grad || !m || !v || !(lr > 0.0f)) return hipErrorInvalidValue;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f / (1.0f - powf(b1, (float)step)), c2 = 1.0f / (1.0f - powf(b2, (float)step));
  hipLaunchKernelGGL(wn_adam_kernel, dim3(wn_grid(count)), dim3(256), 0, stream, count, params, grad, m, v, lr, b1, b2,
                     1e-8f, c1, c2);
  return hipGetLastError();
}

size_t eigen_GPU_batch_refine_bufsize(int L, int n)
{
  if (L < 1 || n < 1 || n > 256) return 0;
  return WS_HEAD + RefineLayout(L, n).total;
}

int eigen_GPU_batch_refine_DP(int L, int nm, int n, const double *a, double *x, int ldx, double *w, void *wk,
                              hipStream_t stream, int iters, int *info)
{
  return run_refine(L, nm, n, a, x, ldx, w, wk, stream, iters, info);
}

size_t eigen_GPU_batch_warm_bufsize(int L, int n, int dtype_bytes) { return warm_bytes(L, n, dtype_bytes); }

int eigen_GPU_batch_warm_FP(int L, int nm, int n, int m, float *a, const float *q0, float *w, void *wk,
                            hipStream_t stream, int max_sweeps, int *info, int *sweeps)
{
  return run_warm<float>(L, nm, n, m, a, q0, w, wk, stream, max_sweeps, info, sweeps);
}

int eigen_GPU_batch_warm_DP(int L, int nm, int n, int m, double *a, const double *q0, double *w, void *wk,
                            hipStream_t stream, int max_sweeps, int *info, int *sweeps)
{
  return run_warm<double>(L, nm, n, m, a, q0, w, wk, stream, max_sweeps, info, sweeps);
}

int eigen_GPU_batch_qr_FP(int L, int nm, int n, const float *a, float *q, hipStream_t stream)
{
  return run_qr<float>(L, nm, n, a, q, stream);
}

int eigen_GPU_batch_qr_DP(int L, int nm, int n, const double *a, double *q, hipStream_t stream)
{
  return run_qr<double>(L, nm, n, a, q, stream);
}

size_t eigen_GPU_batch_interleaved_bufsize(int L, int nm, int n, int m, int dtype_bytes, int solver)
{
  return interleaved_bytes(L, nm, n, m, dtype_bytes, solver);
}

int eigen_GPU_batch_interleaved_FP(int L, int nm, int n, int m, float *a, float *w, void *wk, hipStream_t stream,
                                   int solver, int *info)
{
  return run_interleaved<float>(L, nm, n, m, a, w, wk, stream, solver, info);
}

int eigen_GPU_batch_interleaved_DP(int L, int nm, int n, int m, double *a, double *w, void *wk,
                                   hipStream_t stream, int solver, int *info)
{
  return run_interleaved<double>(L, nm, n, m, a, w, wk, stream, solver, info);
}

int eigen_GPU_batch_gemm(int kind, int transa, int transb, int M, int N, int K, double alpha, const void *A,
                         long long lda, long long stride_a, const void *B, long long ldb, long long stride_b,
                         double beta, const void *D, long long ldd, long long stride_d, void *C, long long ldc,
                         long long stride_c, int batch, hipStream_t stream)
{
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K; g.batch = batch;
  g.transa = transa ? 1 : 0;
  g.transb = transb ? 1 : 0;
  g.A = A; g.lda = lda; g.sa = stride_a;
  g.B = B; g.ldb = ldb; g.sb = stride_b;
  g.D = D; g.ldd = D ? ldd : 0; g.sd = D ? stride_d : 0;
  g.C = C; g.ldc = ldc; g.sc = stride_c;
  g.alpha = alpha;
  g.beta = D ? beta : 0.0;
  g.sym = 0;
  return gemm_launch(kind, g, stream);
}

const char *eigen_GPU_batch_version(void) { return "eigen-mi355x 0.3"; }

} // extern "C"

// the no-allocation entry points (scratch array in device memory)
#include "eig_noalloc.hpp"
