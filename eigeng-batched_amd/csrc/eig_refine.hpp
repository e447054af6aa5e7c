// eig_refine.hpp - FP64 refinement of approximate eigenpairs on the GPU
// (Ogita-Aishima iteration with Rayleigh-Ritz on clusters), the second half of
// the mixed-precision solver: an FP32 solve gives X, then per iteration
//
//   P = A X,  S = X^T P,  G = X^T X         (FP64 products on MFMA, eig_gemm.hpp)
//   R = I - G,  l_i = s_ii / g_ii,
//   delta = 2 (||S - diag(l)||_F + max|l| ||R||_F)
//   E_ij = (s_ij - l_j g_ij) / (l_j - l_i)  (|l_i - l_j| > delta),  E_ii = r_ii / 2
//   X <- X + X E                            (FP64 product on MFMA)
//
// (Ogita & Aishima, "Iterative refinement for symmetric eigenvalue
// decomposition", Japan J. Indust. Appl. Math. 35 (2018), algorithm 1; the
// reference lists "Half precision version + Iterative refinment" as future
// work, README:169.)
//
// Clusters (runs of ascending l with gaps <= delta, where the first-order
// formula does not apply) of up to RF_KMAX members are resolved by
// Rayleigh-Ritz, as in the clustered-eigenvalue variant of the same authors
// (2019): the cluster's basis corrected to second order, orthonormalised and
// rotated onto the eigenvectors of its projected matrix (rf_resolve).  A
// larger cluster keeps E_ij = r_ij / 2 and marks the matrix for the plain FP64
// solver.
//
// Adjacent eigenvalues also join one cluster when the first-order correction
// between them would exceed RF_TAU (|s_ab - l g_ab| > RF_TAU |l_b - l_a|): the
// first-order update leaves an error of the order of E^2, so large entries are
// better resolved exactly.  A matrix is accepted when the last update's largest
// first-order entry (E outside the resolved clusters' own blocks) is at most
// `accept`, so that what the update leaves (~E^2) is at FP64 rounding level;
// otherwise (or on a non-finite value, a cluster beyond RF_KMAX, or a failed
// FP32 solve) its status is set to ST_OVERFLOW and the plain FP64 workgroup
// solver (eig_global_kernel with `only`) solves it from the input, which the
// store kernel leaves untouched for such matrices.
#pragma once
#include "common.hpp"
#include "eig_wave.hpp"

namespace eigmi {

constexpr int RF_THREADS = 256;
constexpr int RF_KMAX = 8;   // largest cluster resolved by Rayleigh-Ritz
constexpr int RF_PAR = 32;   // clusters resolved at once (one thread each)
constexpr double RF_TAU = 1e-2;  // first-order entries beyond this are resolved by Rayleigh-Ritz
// reasons recorded next to the status (diagnostics, tests)
constexpr int RF_WHY_SOLVE = 1, RF_WHY_CLUSTER = 2, RF_WHY_CONV = 4, RF_WHY_NAN = 8;

// ---- symmetric full copy from the upper triangle, with a type change ----------
// One 64 x 64 tile (I, J), I <= J, of one matrix per workgroup: the tile is read
// with rows contiguous and written both in place and, through LDS, transposed
// to (J, I); on the diagonal tile the lower half takes the mirrored upper half.
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void eig_symfull_kernel(int L, int n, const TI *__restrict__ in, long long ldi,
                                                          long long si, TO *__restrict__ out, long long ldo,
                                                          long long so)
{
  __shared__ TI T[64][65];
  const int nt = (n + 63) >> 6, per = nt * (nt + 1) / 2;
  const long long total = (long long)L * per;
  const int rr = threadIdx.x & 63, cq = threadIdx.x >> 6;
  for (long long bid = blockIdx.x; bid < total; bid += gridDim.x) {
    const long long b = bid / per;
    const int r = (int)(bid - b * per);
    int J = 0;
    while ((J + 1) * (J + 2) / 2 <= r) ++J;
    const int I = r - J * (J + 1) / 2;
    const TI *__restrict__ A = in + b * si;
    TO *__restrict__ O = out + b * so;
    const int r0 = I * 64, c0 = J * 64;
#pragma unroll 4
    for (int k = 0; k < 16; ++k) {
      const int cc = cq + 4 * k, i = r0 + rr, j = c0 + cc;
      T[rr][cc] = (i < n && j < n) ? A[i + (size_t)j * ldi] : TI(0);
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < 16; ++k) {
      const int cc = cq + 4 * k;
      const int i = r0 + rr, j = c0 + cc;
      if (i < n && j < n) O[i + (size_t)j * ldo] = (TO)((I < J || rr <= cc) ? T[rr][cc] : T[cc][rr]);
      if (I < J) {
        const int i2 = c0 + rr, j2 = r0 + cc;
        if (i2 < n && j2 < n) O[i2 + (size_t)j2 * ldo] = (TO)T[cc][rr];
      }
    }
    __syncthreads();
  }
}

// y = (TO) x, count elements
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void eig_convert_kernel(long long count, const TI *__restrict__ x, TO *__restrict__ y)
{
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < count; i += (long long)gridDim.x * 256)
    y[i] = (TO)x[i];
}

// n columns of n rows from a source with leading dimension ld into a packed
// (ld n) array, per matrix of a chunk
template <typename T>
__global__ __launch_bounds__(256) void eig_copy_cols_kernel(int L, int n, const T *__restrict__ src, long long ld,
                                                            T *__restrict__ dst)
{
  for (long long c = blockIdx.x; c < (long long)L * n; c += gridDim.x) {
    const long long b = c / n, j = c - b * n;
    const T *__restrict__ sc = src + b * ld * n + j * ld;
    T *__restrict__ dc = dst + (b * n + j) * n;
    for (int r = threadIdx.x; r < n; r += 256) dc[r] = sc[r];
  }
}

// info of the matrices a refinement did not accept: the reasons
__global__ __launch_bounds__(256) void eig_refine_info_kernel(int L, const int *__restrict__ status,
                                                              const int *__restrict__ why, int *__restrict__ info)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < L && status[i] != 0) info[i] = why[i];
}

// status / reasons of a refinement chunk from the FP32 solve's status (none:
// every matrix starts accepted)
__global__ __launch_bounds__(256) void eig_refine_init_kernel(int L, const int *__restrict__ info32,
                                                              int *__restrict__ status, int *__restrict__ why)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < L) {
    const bool bad = info32 && info32[i] != 0;
    status[i] = bad ? ST_OVERFLOW : 0;
    why[i] = bad ? RF_WHY_SOLVE : 0;
  }
}

// Per-matrix results of the statistics step, read by the fill step: old and
// refined eigenvalue estimates, ascending order, cluster of each index (first
// sorted position), position inside it and its size, and per resolved cluster
// starting at sorted position p the k x k matrices C W (at mw + p KMAX) and W
// (at mw + (n + p) KMAX).
struct RfInfo {
  double *lam, *lnew, *mw;
  int *ord, *cl, *pos, *csz;
};
__host__ __device__ inline size_t rf_info_bytes(int n)
{
  return (size_t)n * (2 * 8 + 4 * 4) + (size_t)2 * n * RF_KMAX * 8;
}
// the arrays of matrix b inside a region of lc matrices
__host__ __device__ inline RfInfo rf_info_at(char *base, int lc, int n, long long b)
{
  RfInfo r;
  const size_t nl = (size_t)lc * n;
  r.lam = (double *)base + (size_t)b * n;
  r.lnew = (double *)base + nl + (size_t)b * n;
  r.mw = (double *)base + 2 * nl + (size_t)b * 2 * n * RF_KMAX;
  int *ib = (int *)((double *)base + 2 * nl + nl * 2 * RF_KMAX);
  r.ord = ib + (size_t)b * n;
  r.cl = ib + nl + (size_t)b * n;
  r.pos = ib + 2 * nl + (size_t)b * n;
  r.csz = ib + 3 * nl + (size_t)b * n;
  return r;
}

struct RfLds {
  double lam[256];
  int ord[256];
  int csz[256];
  int mlist[128];
  unsigned char jn[256];                        // order positions q joined to q - 1
  double slab[RF_KMAX * RF_KMAX];               // the cluster's k x k scratch
  double cs[RF_KMAX * RF_KMAX], cg[RF_KMAX * RF_KMAX];   // second-order sums of S', G'
  double pr[RF_THREADS / 64][16];               // per-wave partial sums of a batch
  double red[3][RF_THREADS / 64];
  int nres, why;
};

__device__ __forceinline__ double rf_wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double rf_wave_max(double v)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// element (i, j) of a symmetric matrix whose upper triangle is stored
__device__ __forceinline__ double rf_sym(const double *__restrict__ X, int n, int i, int j)
{
  return X[(size_t)min(i, j) + (size_t)max(i, j) * n];   // the upper triangle; one load, no branch
}

// Rayleigh-Ritz on the cluster J at sorted positions p0 .. p0+k-1 (one thread).
// The cluster's basis is first corrected to second order: Y_J = X_J C + X_K F
// (K the indices outside J, C = I + R_JJ / 2, F the first-order entries
// F_ia = (s_ia - l_a g_ia) / (l_a - l_i)), whose projected matrices are, with
// S_KK ~ diag(l_K) and G_KK ~ I (errors of third order),
//   S' = C^T S_JJ C + S_JK F + F^T S_KJ + F^T diag(l_K) F
//   G' = C^T G_JJ C + G_JK F + F^T G_KJ + F^T F;
// then P = I - H/2 + 3/8 H^2 ~ G'^(-1/2) (H = G' - I), the ascending
// eigenvectors W of P S' P by cyclic Jacobi (FP64), and X'_J = Y_J P W:
// Mc = C P W (the cluster's own block of I + E), Wc = P W (F P W is the block
// outside).  Without the correction the projected matrix would carry the
// subspace's second-order error, which against the cluster's own tiny gaps
// would leave the rotation inside the cluster inaccurate.  sp: k x k scratch.
__device__ void rf_resolve(const int *ord, const double *cs, const double *cg, int k, int p0,
                           const double *__restrict__ Sb, const double *__restrict__ Gb, int n,
                           double *__restrict__ sp, double *__restrict__ Mc, double *__restrict__ Wc,
                           double *__restrict__ lnew)
{
  int idx[RF_KMAX];
#pragma unroll
  for (int a = 0; a < RF_KMAX; ++a) idx[a] = (a < k) ? ord[p0 + a] : 0;
  auto cmat = [&](int a, int b) -> double {       // C = I + R_JJ / 2, R = I - G
    const double g = rf_sym(Gb, n, idx[a], idx[b]);
    return (a == b) ? 1.5 - 0.5 * g : -0.5 * g;
  };
  // sp <- C^T S_JJ C, Mc <- C^T G_JJ C (Wc as the intermediate)
  for (int pass = 0; pass < 2; ++pass) {
    const double *X = pass == 0 ? Sb : Gb;
    double *out = pass == 0 ? sp : Mc;
    for (int a = 0; a < k; ++a)
      for (int b = 0; b < k; ++b) {
        double t = 0.0;
        for (int c = 0; c < k; ++c) t = fma(rf_sym(X, n, idx[a], idx[c]), cmat(c, b), t);
        Wc[a * k + b] = t;
      }
    for (int a = 0; a < k; ++a)
      for (int b = 0; b < k; ++b) {
        double t = 0.0;
        for (int c = 0; c < k; ++c) t = fma(cmat(c, a), Wc[c * k + b], t);
        out[a * k + b] = t;
      }
  }
  // the second-order terms over the indices outside J (rf_cluster_sums)
  for (int a = 0; a < k; ++a)
    for (int b = 0; b < k; ++b) {
      sp[a * k + b] += cs[a * RF_KMAX + b];
      Mc[a * k + b] += cg[a * RF_KMAX + b];
    }
  // P = I - H/2 + 3/8 H^2, H = G' - I  (into Wc)
  for (int a = 0; a < k; ++a)
    for (int b = 0; b < k; ++b) {
      double h2 = 0.0;
      for (int c = 0; c < k; ++c)
        h2 = fma(Mc[a * k + c] - (a == c ? 1.0 : 0.0), Mc[c * k + b] - (c == b ? 1.0 : 0.0), h2);
      const double h = Mc[a * k + b] - (a == b ? 1.0 : 0.0);
      Wc[a * k + b] = (a == b ? 1.0 : 0.0) - 0.5 * h + 0.375 * h2;
    }
  // sp <- P S' P  (S' P into Mc first)
  for (int a = 0; a < k; ++a)
    for (int b = 0; b < k; ++b) {
      double t = 0.0;
      for (int c = 0; c < k; ++c) t = fma(sp[a * k + c], Wc[c * k + b], t);
      Mc[a * k + b] = t;
    }
  double fro = 0.0;
  for (int a = 0; a < k; ++a)
    for (int b = 0; b < k; ++b) {
      double t = 0.0;
      for (int c = 0; c < k; ++c) t = fma(Wc[a * k + c], Mc[c * k + b], t);
      sp[a * k + b] = t;
      fro += t * t;
    }
  // W (in Mc) by cyclic Jacobi, FP64 (Golub & Van Loan, symmetric Schur 2 x 2)
  for (int a = 0; a < k; ++a)
    for (int b = 0; b < k; ++b) Mc[a * k + b] = (a == b) ? 1.0 : 0.0;
  const double tol2 = fro * 4.93e-32;             // (eps)^2 ||sp||_F^2
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0;
    for (int a = 0; a < k; ++a)
      for (int b = a + 1; b < k; ++b) off += sp[a * k + b] * sp[a * k + b];
    if (off <= tol2) break;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        const double apq = sp[p * k + q];
        if (apq == 0.0) continue;
        const double app = sp[p * k + p], aqq = sp[q * k + q];
        const double th = (aqq - app) / (2.0 * apq);
        const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(fma(th, th, 1.0)));
        const double c = 1.0 / sqrt(fma(t, t, 1.0)), s = t * c;
        for (int r = 0; r < k; ++r) {             // columns p, q
          const double xp = sp[r * k + p], xq = sp[r * k + q];
          sp[r * k + p] = c * xp - s * xq;
          sp[r * k + q] = s * xp + c * xq;
        }
        for (int r = 0; r < k; ++r) {             // rows p, q
          const double xp = sp[p * k + r], xq = sp[q * k + r];
          sp[p * k + r] = c * xp - s * xq;
          sp[q * k + r] = s * xp + c * xq;
        }
        for (int r = 0; r < k; ++r) {
          const double xp = Mc[r * k + p], xq = Mc[r * k + q];
          Mc[r * k + p] = c * xp - s * xq;
          Mc[r * k + q] = s * xp + c * xq;
        }
      }
  }
  // ascending eigenvalues, W's columns with them (selection sort, first minimum)
  for (int a = 0; a < k; ++a) {
    int mi = a;
    for (int b = a + 1; b < k; ++b)
      if (sp[b * k + b] < sp[mi * k + mi]) mi = b;
    if (mi != a) {
      const double t = sp[a * k + a];
      sp[a * k + a] = sp[mi * k + mi];
      sp[mi * k + mi] = t;
      for (int r = 0; r < k; ++r) {
        const double x = Mc[r * k + a];
        Mc[r * k + a] = Mc[r * k + mi];
        Mc[r * k + mi] = x;
      }
    }
  }
  // the member at cluster position a takes eigenvalue a
  for (int a = 0; a < k; ++a) lnew[idx[a]] = sp[a * k + a];
  // sp <- P W, Wc <- P W, Mc <- C P W
  for (int a = 0; a < k; ++a)
    for (int b = 0; b < k; ++b) {
      double t = 0.0;
      for (int c = 0; c < k; ++c) t = fma(Wc[a * k + c], Mc[c * k + b], t);
      sp[a * k + b] = t;
    }
  for (int a = 0; a < k * k; ++a) Wc[a] = sp[a];
  for (int a = 0; a < k; ++a)
    for (int b = 0; b < k; ++b) {
      double t = 0.0;
      for (int c = 0; c < k; ++c) t = fma(cmat(a, c), sp[c * k + b], t);
      Mc[a * k + b] = t;
    }
}

// (a, b), a <= b, of pair number q in row order of a K x K upper triangle
__host__ __device__ constexpr int rf_pair_a(int q, int K)
{
  int a = 0;
  while (q >= K - a) { q -= K - a; ++a; }
  return a;
}
__host__ __device__ constexpr int rf_pair_b(int q, int K)
{
  int a = 0;
  while (q >= K - a) { q -= K - a; ++a; }
  return a + q;
}

// The second-order sums of rf_resolve for the cluster of K members at sorted
// position p0, over the indices outside it, one index per thread (n <= 256):
//   cs_ab = sum_i s_ia f_ib + f_ia s_ib + l_i f_ia f_ib,
//   cg_ab = sum_i g_ia f_ib + f_ia g_ib + f_ia f_ib,
// f_ia = (s_ia - l_a g_ia) / (l_a - l_i); pairs reduced over the workgroup in
// batches of 8 (16 values), both halves of cs / cg written.  All threads call.
template <int K>
__device__ void rf_cluster_sums(RfLds &sh, int p0, const double *__restrict__ Sb, const double *__restrict__ Gb,
                                int n)
{
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  int idx[K];
#pragma unroll
  for (int a = 0; a < K; ++a) idx[a] = sh.ord[p0 + a];
  bool use = t < n;
#pragma unroll
  for (int a = 0; a < K; ++a) use = use && idx[a] != t;
  const double li = use ? sh.lam[t] : 0.0;
  // every thread loads (row 0 beyond n) and masks: the 2 K loads are in flight together
  const int tc = t < n ? t : 0;
  double si[K], gi[K], f[K];
#pragma unroll
  for (int a = 0; a < K; ++a) {
    const double sv = rf_sym(Sb, n, tc, idx[a]), gv = rf_sym(Gb, n, tc, idx[a]);
    si[a] = use ? sv : 0.0;
    gi[a] = use ? gv : 0.0;
    const double la = sh.lam[idx[a]];
    f[a] = use ? (si[a] - la * gi[a]) / (la - li) : 0.0;
  }
  constexpr int NP = K * (K + 1) / 2;
#pragma unroll
  for (int q0 = 0; q0 < NP; q0 += 8) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int q = q0 + u;
      double vs = 0.0, vg = 0.0;
      if (q < NP) {
        const int a = rf_pair_a(q, K), b = rf_pair_b(q, K);
        vs = si[a] * f[b] + f[a] * si[b] + li * f[a] * f[b];
        vg = gi[a] * f[b] + f[a] * gi[b] + f[a] * f[b];
      }
      v[2 * u] = rf_wave_sum(vs);
      v[2 * u + 1] = rf_wave_sum(vg);
    }
    if (lane == 0)
#pragma unroll
      for (int u = 0; u < 16; ++u) sh.pr[wv][u] = v[u];
    __syncthreads();
    if (t < 16 && q0 + (t >> 1) < NP) {
      double sum = 0.0;
#pragma unroll
      for (int w = 0; w < RF_THREADS / 64; ++w) sum += sh.pr[w][t];
      const int q = q0 + (t >> 1);
      const int a = rf_pair_a(q, K), b = rf_pair_b(q, K);
      double *dst = (t & 1) ? sh.cg : sh.cs;
      dst[a * RF_KMAX + b] = sum;
      dst[b * RF_KMAX + a] = sum;
    }
    __syncthreads();
  }
}

// Statistics step of one refinement iteration, one workgroup per matrix
// (n <= 256): Rayleigh quotients, the norms and delta, ascending order,
// clusters, Rayleigh-Ritz on the clusters of 2..KMAX members.  S = X^T A X and
// G = X^T X hold their upper triangles (ld n).  status / why: see the file
// comment (set, never cleared, so a matrix that failed earlier stays marked).
__global__ __launch_bounds__(RF_THREADS) void eig_oa_stats_kernel(int L, int n, const double *__restrict__ S,
                                                                  const double *__restrict__ G, char *__restrict__ inf,
                                                                  int lc, int *__restrict__ status,
                                                                  int *__restrict__ why,
                                                                  unsigned long long *__restrict__ emax)
{
  __shared__ RfLds sh;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const size_t nn = (size_t)n * n;
  for (long long b = blockIdx.x; b < L; b += gridDim.x) {
    const double *__restrict__ Sb = S + b * nn;
    const double *__restrict__ Gb = G + b * nn;
    const RfInfo ri = rf_info_at(inf, lc, n, b);
    double so = 0.0, go = 0.0, mx = 0.0;
    bool bad = false;
    if (t < n) {
      const double s = Sb[t + (size_t)t * n], g = Gb[t + (size_t)t * n];
      const double l = s / g;
      sh.lam[t] = l;
      ri.lam[t] = l;
      ri.lnew[t] = l;
      so = (s - l) * (s - l);
      go = (1.0 - g) * (1.0 - g);
      mx = fabs(l);
      bad = !isfinite(l);
    }
    // strict upper triangle: column j, rows t < j (coalesced, independent
    // loads).  Every thread loads (row 0 where t >= j) and masks the value, so
    // that no branch separates the loads of the unrolled columns and they are
    // all in flight before the first is used
#pragma unroll 16
    for (int j = 1; j < n; ++j) {
      const bool in = t < j;
      const size_t o = (size_t)(in ? t : 0) + (size_t)j * n;
      const double s = Sb[o], g = Gb[o];
      so = fma(2.0 * s, in ? s : 0.0, so);
      go = fma(2.0 * g, in ? g : 0.0, go);
    }
    so = rf_wave_sum(so);
    go = rf_wave_sum(go);
    mx = rf_wave_max(mx);
    const int anybad = __syncthreads_or(bad ? 1 : 0);
    if (lane == 0) { sh.red[0][wv] = so; sh.red[1][wv] = go; sh.red[2][wv] = mx; }
    __syncthreads();
    so = go = mx = 0.0;
#pragma unroll
    for (int w = 0; w < RF_THREADS / 64; ++w) { so += sh.red[0][w]; go += sh.red[1][w]; mx = fmax(mx, sh.red[2][w]); }
    const double nso = sqrt(so), ngo = sqrt(go);
    const double delta = 2.0 * (nso + mx * ngo);
    if (t == 0) emax[b] = 0ull;                      // the fill step's largest first-order entry
    // ascending order (ties by index)
    if (t < n) {
      const double l = sh.lam[t];
      int r = 0;
      for (int j = 0; j < n; ++j) {
        const double lj = sh.lam[j];
        r += (lj < l) || (lj == l && j < t);
      }
      sh.ord[r] = t;
      ri.ord[r] = t;
    }
    __syncthreads();
    // adjacent indices of the order join when their gap is within delta or
    // their first-order correction would exceed RF_TAU: one pair per thread,
    // so that the entries of S and G it may read load in parallel
    if (t >= 1 && t < n) {
      const int ia = sh.ord[t - 1], ib = sh.ord[t];
      const double la = sh.lam[ia], lb = sh.lam[ib], gap = lb - la;
      bool jn = gap <= delta;
      if (!jn) {
        const double s = rf_sym(Sb, n, ia, ib), g = rf_sym(Gb, n, ia, ib);
        jn = fmax(fabs(s - lb * g), fabs(s - la * g)) > RF_TAU * gap;
      }
      sh.jn[t] = jn ? 1 : 0;
    }
    __syncthreads();
    if (t == 0) {
      int why_b = anybad ? RF_WHY_NAN : 0, nres = 0;
      int p = 0;
      while (p < n) {
        int e = p + 1;
        while (e < n && sh.jn[e]) ++e;
        const int k = e - p;
        for (int q = p; q < e; ++q) {
          const int i = sh.ord[q];
          ri.cl[i] = p;
          ri.pos[i] = q - p;
          ri.csz[i] = k;
        }
        sh.csz[p] = k;
        if (k >= 2) {
          if (k <= RF_KMAX) sh.mlist[nres++] = p;
          else why_b |= RF_WHY_CLUSTER;
        }
        p = e;
      }
      sh.nres = nres;
      sh.why = why_b;
    }
    __syncthreads();
    // clusters one at a time: the second-order sums over the indices outside
    // the cluster with one index per thread (rf_cluster_sums), then the small
    // k x k work on thread 0
    for (int c = 0; c < sh.nres; ++c) {
      const int p0 = sh.mlist[c], k = sh.csz[p0];
      switch (k) {
      case 2: rf_cluster_sums<2>(sh, p0, Sb, Gb, n); break;
      case 3: rf_cluster_sums<3>(sh, p0, Sb, Gb, n); break;
      case 4: rf_cluster_sums<4>(sh, p0, Sb, Gb, n); break;
      case 5: rf_cluster_sums<5>(sh, p0, Sb, Gb, n); break;
      case 6: rf_cluster_sums<6>(sh, p0, Sb, Gb, n); break;
      case 7: rf_cluster_sums<7>(sh, p0, Sb, Gb, n); break;
      default: rf_cluster_sums<8>(sh, p0, Sb, Gb, n); break;
      }
      if (t == 0)
        rf_resolve(sh.ord, sh.cs, sh.cg, k, p0, Sb, Gb, n, sh.slab, ri.mw + (size_t)p0 * RF_KMAX,
                   ri.mw + (size_t)(n + p0) * RF_KMAX, ri.lnew);
      __syncthreads();
    }
    if (t == 0 && sh.why) {
      why[b] |= sh.why;
      status[b] = ST_OVERFLOW;
    }
    __syncthreads();
  }
}

// Fill step: E for one 64 x 64 tile (I, J) of one matrix per workgroup.  The
// stored (upper) tile of S and G is staged in LDS, so the lower tiles read it
// transposed and every global access is coalesced.
__global__ __launch_bounds__(256) void eig_oa_fill_kernel(int L, int n, const double *__restrict__ S,
                                                          const double *__restrict__ G, char *__restrict__ inf,
                                                          int lc, double *__restrict__ E,
                                                          unsigned long long *__restrict__ emax)
{
  __shared__ double wmx[4];
  __shared__ double Ts[64][65], Tg[64][65];
  __shared__ double li[64], lj[64];
  __shared__ int ci[64], cj[64], kj[64], pj[64], pi_[64];
  const int nt = (n + 63) >> 6, per = nt * nt;
  const long long total = (long long)L * per;
  const int rr = threadIdx.x & 63, cq = threadIdx.x >> 6;
  const size_t nn = (size_t)n * n;
  for (long long bid = blockIdx.x; bid < total; bid += gridDim.x) {
    const long long b = bid / per;
    const int tix = (int)(bid - b * per);
    const int I = tix % nt, J = tix / nt;
    const bool up = I <= J;
    const int lr0 = (up ? I : J) * 64, lc0 = (up ? J : I) * 64;   // the stored tile
    const double *__restrict__ Sb = S + b * nn;
    const double *__restrict__ Gb = G + b * nn;
    const RfInfo ri = rf_info_at(inf, lc, n, b);
    // every thread loads (indices clamped into the matrix) and masks the
    // value: no branch between the loads, so they are all in flight at once
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int cc = cq + 4 * k, i = lr0 + rr, j = lc0 + cc;
      const bool in = i < n && j < n;
      const size_t o = (size_t)min(i, n - 1) + (size_t)min(j, n - 1) * n;
      const double sv = Sb[o], gv = Gb[o];
      Ts[rr][cc] = in ? sv : 0.0;
      Tg[rr][cc] = in ? gv : 0.0;
    }
    if (threadIdx.x < 64) {
      const int i = I * 64 + rr;
      li[rr] = i < n ? ri.lam[i] : 0.0;
      ci[rr] = i < n ? ri.cl[i] : -1;
      pi_[rr] = i < n ? ri.pos[i] : 0;
    } else if (threadIdx.x < 128) {
      const int j = J * 64 + rr;
      lj[rr] = j < n ? ri.lam[j] : 0.0;
      cj[rr] = j < n ? ri.cl[j] : -2;
      kj[rr] = j < n ? ri.csz[j] : 1;
      pj[rr] = j < n ? ri.pos[j] : 0;
    }
    __syncthreads();
    double *__restrict__ Eb = E + b * nn;
    double fo = 0.0;                                 // largest first-order entry of this thread
#pragma unroll 2
    for (int k = 0; k < 16; ++k) {
      const int cc = cq + 4 * k;
      const int i = I * 64 + rr, j = J * 64 + cc;
      if (i < n && j < n) {
        // (i, j) inside the stored tile: (rr, cc) for upper tiles, (cc, rr) below;
        // on the diagonal tile the lower half mirrors the upper half
        const bool direct = (I < J) || (I == J && rr <= cc);
        const double s = direct ? Ts[rr][cc] : Ts[cc][rr];
        const double g = direct ? Tg[rr][cc] : Tg[cc][rr];
        const int kc = kj[cc];
        const bool same = ci[rr] == cj[cc];
        double e;
        if (kc < 2 || kc > RF_KMAX) {
          if (i == j) e = 0.5 * (1.0 - g);
          else if (same) e = -0.5 * g;
          else e = (s - lj[cc] * g) / (lj[cc] - li[rr]);
          fo = fmax(fo, fabs(e));
        } else {
          const int a = pj[cc], c0 = cj[cc];
          if (same) {
            e = ri.mw[(size_t)c0 * RF_KMAX + pi_[rr] * kc + a] - (i == j ? 1.0 : 0.0);
          } else {
            const double *__restrict__ Wc = ri.mw + (size_t)(n + c0) * RF_KMAX;
            e = 0.0;
            for (int c = 0; c < kc; ++c) {
              const int jc = ri.ord[c0 + c];
              const double lc_ = ri.lam[jc];
              const double f = (rf_sym(Sb, n, i, jc) - lc_ * rf_sym(Gb, n, i, jc)) / (lc_ - li[rr]);
              e = fma(f, Wc[c * kc + a], e);
            }
            fo = fmax(fo, fabs(e));
          }
        }
        Eb[i + (size_t)j * n] = e;
      }
    }
    // non-negative doubles order like their bit patterns: one atomic per workgroup
    fo = rf_wave_max(fo);
    if ((threadIdx.x & 63) == 0) wmx[threadIdx.x >> 6] = fo;
    __syncthreads();
    if (threadIdx.x == 0) {
      const double m = fmax(fmax(wmx[0], wmx[1]), fmax(wmx[2], wmx[3]));
      atomicMax(emax + b, (unsigned long long)__double_as_longlong(m));
    }
    __syncthreads();
  }
}

// Final store: eigenvalues ascending into w (stride n), the eigenvectors in the
// same order into a (leading dimension nm; only the first m columns), info 0;
// matrices not accepted are marked for the fallback and left untouched (a
// still holds the input) unless write_all.  sorted == 0 keeps the refined
// order (the FP32 solve's).
__global__ __launch_bounds__(RF_THREADS) void eig_refine_store_kernel(int L, int nm, int n, int m,
                                                                      const double *__restrict__ X,
                                                                      const double *__restrict__ lam,
                                                                      int *__restrict__ status,
                                                                      int *__restrict__ why,
                                                                      const unsigned long long *__restrict__ emax,
                                                                      double accept, double *__restrict__ a,
                                                                      double *__restrict__ w,
                                                                      int *__restrict__ info, int sorted,
                                                                      int write_all)
{
  __shared__ int src[GL_NMAX];
  __shared__ double ls[GL_NMAX];
  const int t = threadIdx.x;
  const size_t nn = (size_t)n * n;
  for (long long b = blockIdx.x; b < L; b += gridDim.x) {
    // acceptance: the last update's first-order entries at most `accept`
    const double em = __longlong_as_double((long long)emax[b]);
    const bool reject = status[b] != 0 || !(em <= accept);   // uniform over the workgroup
    __syncthreads();
    if (reject) {
      if (t == 0 && status[b] == 0) {
        status[b] = ST_OVERFLOW;
        why[b] |= RF_WHY_CONV;
      }
      if (!write_all) continue;
    }
    for (int i = t; i < n; i += RF_THREADS) ls[i] = lam[b * n + i];
    __syncthreads();
    for (int i = t; i < n; i += RF_THREADS) {
      int r = i;
      if (sorted) {
        const double l = ls[i];
        r = 0;
        for (int j = 0; j < n; ++j) r += (ls[j] < l) || (ls[j] == l && j < i);
      }
      src[r] = i;
      w[b * n + r] = ls[i];
    }
    __syncthreads();
    const double *__restrict__ Xb = X + b * nn;
    double *__restrict__ Ab = a + b * (size_t)nm * n;
    for (int j = 0; j < m; ++j) {
      const int s = src[j];
      for (int r = t; r < n; r += RF_THREADS) Ab[r + (size_t)j * nm] = Xb[r + (size_t)s * n];
    }
    if (t == 0 && info && !reject) info[b] = 0;
    __syncthreads();
  }
}

} // namespace eigmi
