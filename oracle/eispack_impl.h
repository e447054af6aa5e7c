/*
 * CPU oracle for the batched symmetric eigensolver: TEST INFRASTRUCTURE ONLY.
 *
 * This file is a C restatement of the algorithms the reference library runs per
 * matrix.  It is the checker the GPU library is compared against, and the CPU
 * baseline that the benchmark reports beside the GPU number.  It is never linked
 * into, called by, or used as a fallback for the GPU library.
 *
 * It is included twice by eispack.c, once with REAL=float and once with
 * REAL=double, so every routine exists "in type" (the arithmetic the FP32 / FP64
 * API computes in).  All matrices are column-major with leading dimension nm;
 * the comments use 1-based (Fortran/EISPACK) indices like the reference.
 *
 * Reference conventions followed (file:line in RIKEN-RCCS/EigenG-Batched):
 *   tridiagonalisation   tred1_tiled.hpp:42-212, hhsy2tr.hpp:57-313
 *   implicit-shift QL    imtql2_tiled.hpp:43-164 (EISPACK imtql2)
 *   explicit-shift QL    tql2_tiled.hpp:44-181   (EISPACK tql2)
 *   ascending sort       tql2_tiled.hpp:149-181
 *   back-transformation  hhtr2sy_tiled.hpp:45-221, hhtr2sy.hpp:46-215
 *   Jacobi               jac_tiled.hpp:20-147, jac.cpp:60-217
 *   accuracy metrics     eigen_GPU_check.cu:34-81
 *   hypot                misc_gpu.hpp:514-548 (Borges 2020, compensated)
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)
#define A_(r, c) a[(size_t)((r)-1) + (size_t)((c)-1) * (size_t)nm]
#define Z_(r, c) z[(size_t)((r)-1) + (size_t)((c)-1) * (size_t)nm]

/* |x| with the sign of y (copysign), as Sign() in misc_gpu.hpp:441-469. */
static inline REAL FN(sgn)(REAL x, REAL y) { return (REAL)CAT(copysign, MSFX)(x, y); }
static inline REAL FN(fabs_)(REAL x) { return (REAL)CAT(fabs, MSFX)(x); }
static inline REAL FN(sqrt_)(REAL x) { return (REAL)CAT(sqrt, MSFX)(x); }
static inline REAL FN(fma_)(REAL x, REAL y, REAL z) { return (REAL)CAT(fma, MSFX)(x, y, z); }
static inline REAL FN(nz)(REAL x) { return x == (REAL)0 ? (REAL)1 : x; } /* flip0to1 */

/* sqrt(x*x + y*y) with one compensated Newton correction (Borges, ACM TOMS
 * 2020, "An improved algorithm for hypot(a,b)"): the rounding errors of the
 * squares and of the sum are carried in tau and folded back in.  Follows the
 * method of misc_gpu.hpp:533-548; no scaling (the reference does none). */
static REAL FN(pythag)(REAL x, REAL y)
{
  const REAL xx = x * x, yy = y * y;
  const REAL sig = xx + yy;
  const REAL h = FN(sqrt_)(sig);
  const REAL sig_err = (sig - xx) - yy;
  const REAL tau = FN(fma_)(y, y, -yy) + FN(fma_)(x, x, -xx) - sig_err + FN(fma_)(-h, h, sig);
  return FN(fma_)(tau / FN(nz)(h), (REAL)0.5, h);
}
static REAL FN(pythag1)(REAL x) { return FN(pythag)(x, (REAL)1); }

/* Householder tridiagonalisation, reference storage convention
 * (tred1_tiled.hpp:42-212):
 *   reads only the upper triangle a(j,k), j<=k;
 *   for i = n..2 (l = i-1): u = a(1:l,i), anorm = -sign(||u||, u_l),
 *     u_l <- u_l - anorm, beta = 1/(u_l*anorm) (1 if that product is 0),
 *     p = beta*A11*u, v = p + (beta/2)(u.p) u, A11 += u v^T + v u^T,
 *     e(l) = anorm and u stays in a(1:l,i);
 *   then d(j) = a(j,j), a(j,j) <- e(j-1) for j>=2, a(1,1) <- e(n) = 0.
 * H_i = I + beta u u^T, and T = H_2 ... H_n A H_n ... H_2.
 * Unlike tred1_tiled.hpp:26-34, n=2 runs the general step (a 1x1 reflector),
 * so the back-transformation stays consistent for n=2. */
static void FN(tred1)(int nm, int n, REAL *a, REAL *d, REAL *e, REAL *work)
{
  REAL *u = work, *v = work + n;
  for (int i = n; i >= 2; --i) {
    const int l = i - 1;
    REAL s = 0;
    for (int k = 1; k <= l; ++k) { u[k] = A_(k, i); s += u[k] * u[k]; }
    const REAL anorm = -FN(sgn)(FN(sqrt_)(s), u[l]);
    const REAL ul = u[l] - anorm;
    const REAL beta = (REAL)1 / FN(nz)(ul * anorm);
    u[l] = ul;
    A_(l, i) = ul;
    e[l - 1] = anorm;
    for (int j = 1; j <= l; ++j) v[j] = 0;
    for (int k = 1; k <= l; ++k) {       /* p = A11 u from the upper triangle */
      for (int j = 1; j < k; ++j) {
        const REAL ajk = A_(j, k);
        v[j] += ajk * u[k];
        v[k] += ajk * u[j];
      }
      v[k] += A_(k, k) * u[k];
    }
    REAL alpha = 0;
    for (int j = 1; j <= l; ++j) { v[j] *= beta; alpha += v[j] * u[j]; }
    alpha *= beta * (REAL)0.5;
    for (int j = 1; j <= l; ++j) v[j] += alpha * u[j];
    for (int k = 1; k <= l; ++k)
      for (int j = 1; j <= k; ++j) A_(j, k) += u[j] * v[k] + v[j] * u[k];
  }
  for (int j = 1; j <= n; ++j) d[j - 1] = A_(j, j);
  e[n - 1] = 0;
  for (int j = 2; j <= n; ++j) A_(j, j) = e[j - 2];
  A_(1, 1) = 0;
}

static void FN(zero_identity)(int nm, int n, REAL *z)
{
  for (int c = 1; c <= n; ++c)
    for (int r = 1; r <= n; ++r) Z_(r, c) = (r == c) ? (REAL)1 : (REAL)0;
}

/* Givens rotation of columns i, i+1 of Z from the right (EISPACK "form vector"). */
static inline void FN(rot)(int nm, int n, REAL *z, int i, REAL c, REAL s)
{
  for (int k = 1; k <= n; ++k) {
    const REAL h0 = Z_(k, i), h1 = Z_(k, i + 1);
    Z_(k, i + 1) = s * h0 + c * h1;
    Z_(k, i) = c * h0 - s * h1;
  }
}

/* Implicit-shift QL with eigenvector accumulation into z (z = I on entry is set
 * here), reference variant of EISPACK imtql2 (imtql2_tiled.hpp:43-164):
 *   deflation |e(m)| <= tol*(|d(m)|+|d(m+1)|), m in [l, n-1], else m = n;
 *   at most max_sweep+1 sweeps per eigenvalue, then returns l (ierror);
 *   Wilkinson shift from the leading 2x2, d(i+1) = fma(s, q, d(i+1)-p).
 * d(1..n) diagonal, e(1..n) with e(l) coupling l and l+1 and e(n) = 0. */
static int FN(imtql2)(int nm, int n, REAL *d_, REAL *e_, REAL *z, REAL tol, int max_sweep)
{
  REAL *d = d_ - 1, *e = e_ - 1;
  FN(zero_identity)(nm, n, z);
  for (int l = 1; l <= n; ++l) {
    int itr = 0;
    for (;;) {
      if (itr > max_sweep) return l;
      int m = n;
      if (l < n)
        for (int mm = l; mm <= n - 1; ++mm)
          if (FN(fabs_)(e[mm]) <= (FN(fabs_)(d[mm]) + FN(fabs_)(d[mm + 1])) * tol) { m = mm; break; }
      if (m == l) break;
      ++itr;
      const REAL dl = d[l], el = e[l];
      const REAL f0 = (d[l + 1] - dl) / (el + el);
      const REAL g0 = FN(pythag1)(f0);
      REAL di1 = d[m];
      REAL r = (di1 - dl) + el / (f0 + FN(sgn)(g0, f0));
      REAL s = 1, c = 1, delta = 0;
      int i;
      for (i = m - 1; i >= l; --i) {
        const REAL ei = e[i];
        const REAL f = s * ei, g = r;
        r = FN(pythag)(f, g);
        if (r == (REAL)0) break;               /* underflow: restart the sweep */
        const REAL b = c * ei;
        s = f / r;
        c = g / r;
        const REAL dix = di1 - delta;
        di1 = d[i];
        const REAL q = (di1 - dix) * s + (REAL)2 * c * b;
        const REAL dx1 = FN(fma_)(s, q, dix);
        delta = dx1 - dix;
        e[i + 1] = r;
        d[i + 1] = dx1;
        r = c * q - b;
        FN(rot)(nm, n, z, i, c, s);
      }
      const int j = (i + 1 > l) ? i + 1 : l;
      d[j] -= delta;
      e[j] = r;
      e[m] = 0;
    }
  }
  return 0;
}

/* Explicit-shift QL with eigenvector accumulation, reference variant of EISPACK
 * tql2 (tql2_tiled.hpp:44-146): tst1 = max(tst1, tol*(|d(l)|+|e(l)|)), at most
 * max_sweep sweeps per eigenvalue, then returns l and stops. */
static int FN(tql2)(int nm, int n, REAL *d_, REAL *e_, REAL *z, REAL tol, int max_sweep)
{
  REAL *d = d_ - 1, *e = e_ - 1;
  FN(zero_identity)(nm, n, z);
  REAL shift = 0, tst1 = 0;
  for (int l = 1; l <= n; ++l) {
    const REAL h = (FN(fabs_)(d[l]) + FN(fabs_)(e[l])) * tol;
    if (h > tst1) tst1 = h;
    int m = n;
    for (int mm = l; mm <= n; ++mm)
      if (FN(fabs_)(e[mm]) <= tst1) { m = mm; break; }
    if (m != l) {
      int itr;
      for (itr = 0; itr < max_sweep; ++itr) {
        const REAL dl_old = d[l], el = e[l];
        const REAL p0 = (d[l + 1] - dl_old) / (el + el);
        const REAL r0 = FN(pythag1)(p0);
        const REAL psr = p0 + FN(sgn)(r0, p0);
        const REAL dl = el / psr;
        const REAL dl1 = el * psr;
        d[l] = dl;
        d[l + 1] = dl1;
        const REAL delta = dl_old - dl;
        for (int i = l + 2; i <= n; ++i) d[i] -= delta;
        shift += delta;
        REAL c = 1, c2 = 1, c3 = 1, s = 0, s2 = 0;
        REAL p = d[m];
        const REAL el1 = e[l + 1];
        REAL ei = 0;
        for (int i = m - 1; i >= l; --i) {
          c3 = c2; c2 = c; s2 = s;
          ei = e[i];
          const REAL di = d[i];
          const REAL g = c * ei, hh = c * p;
          const REAL r = FN(pythag)(p, ei);
          const REAL ei1 = s * r;
          s = ei / r;
          c = p / r;
          p = c * di - s * g;
          const REAL di1 = hh + s * (c * g + s * di);
          e[i + 1] = ei1;
          d[i + 1] = di1;
          FN(rot)(nm, n, z, i, c, s);
        }
        const REAL r = (-s * s2 * c3 * el1 * ei) / dl1;
        e[l] = s * r;
        d[l] = c * r;
        if (FN(fabs_)(e[l]) <= tst1) break;
      }
      if (itr >= max_sweep) return l;
    }
    d[l] += shift;
  }
  return 0;
}

/* Ascending selection sort of d with the reference's tie rule (first minimum,
 * strict '>'; swaps pos(l) and pos(il)), tql2_tiled.hpp:149-176.  pos is 0-based. */
static void FN(sort_pos)(int n, REAL *d, int *pos)
{
  for (int j = 0; j < n; ++j) pos[j] = j;
  for (int l = 0; l < n - 1; ++l) {
    REAL dl = d[l];
    int il = l;
    for (int j = l + 1; j < n; ++j)
      if (dl > d[j]) { dl = d[j]; il = j; }
    if (il != l) {
      const int t = pos[l]; pos[l] = pos[il]; pos[il] = t;
      d[il] = d[l];
      d[l] = dl;
    }
  }
}

/* Back-transformation z <- H_2 H_3 ... H_n z (reflector i ascending), with
 * beta_i = 1/(a(i,i)*a(l,i)) read back from the tred1 storage
 * (hhtr2sy_tiled.hpp:49-68 unblocked form). */
static void FN(trbak)(int nm, int n, const REAL *a, REAL *z)
{
  for (int i = 2; i <= n; ++i) {
    const int l = i - 1;
    const REAL beta = (REAL)1 / FN(nz)(A_(i, i) * A_(l, i));
    for (int j = 1; j <= n; ++j) {
      REAL s = 0;
      for (int k = 1; k <= l; ++k) s += A_(k, i) * Z_(k, j);
      s *= beta;
      for (int k = 1; k <= l; ++k) Z_(k, j) += s * A_(k, i);
    }
  }
}

/* Two-sided Jacobi with the reference's pivoting (jac_tiled.hpp:20-147): for
 * each column i0, ii = argmax_k |a(k,i0)|, jj = argmax_k |a(k,ii)| (ties go to
 * the larger index, as the butterfly at jac_tiled.hpp:48-50 does), rotate (min, max) unless |a(j,i)| < tol*(|w(i)|+|w(j)|); stop
 * after a pass over i0 with no rotation (or after max_pass passes: returns 1).
 * Eigenvalues unsorted in w, eigenvectors in z.  The full matrix is read. */
static int FN(jacobi)(int nm, int n, REAL *a, REAL *z, REAL *w, REAL tol, int max_pass)
{
  FN(zero_identity)(nm, n, z);
  for (int k = 1; k <= n; ++k) { w[k - 1] = A_(k, k); A_(k, k) = 0; }
  for (int pass = 0; pass < max_pass; ++pass) {
    int stable = 1;
    for (int i0 = 1; i0 <= n; ++i0) {
      int ii = 1; REAL x = -1;
      for (int k = 1; k <= n; ++k) { const REAL t = FN(fabs_)(A_(k, i0)); if (t >= x) { x = t; ii = k; } }
      int jj = 1; REAL y = -1;
      for (int k = 1; k <= n; ++k) { const REAL t = FN(fabs_)(A_(k, ii)); if (t >= y) { y = t; jj = k; } }
      int i = ii < jj ? ii : jj, j = ii < jj ? jj : ii;
      if (i == j) continue;
      REAL aij = A_(j, i);
      const REAL aii = w[i - 1], ajj = w[j - 1];
      if (FN(fabs_)(aij) < (FN(fabs_)(aii) + FN(fabs_)(ajj)) * tol) continue;
      const REAL alpha = 2 * aij, beta = ajj - aii;
      const REAL t = alpha / (beta + FN(sgn)(FN(pythag)(alpha, beta), beta));
      const REAL c = (REAL)1 / FN(pythag1)(t);
      const REAL s = c * t;
      const REAL cc = c * c, ss = s * s, cs = -(c * s);
      aij = ((c - s) * (c + s) * aij) + cs * beta;
      const REAL wi = ((cc * aii) + ss * ajj) + cs * alpha;
      const REAL wj = (aii + ajj) - wi;
      w[i - 1] = wi;
      w[j - 1] = wj;
      for (int k = 1; k <= n; ++k) {
        if (k == i || k == j) continue;
        const REAL aki = A_(k, i), akj = A_(k, j);
        const REAL bki = c * aki - s * akj, bkj = s * aki + c * akj;
        A_(i, k) = bki; A_(j, k) = bkj; A_(k, i) = bki; A_(k, j) = bkj;
      }
      A_(i, j) = aij; A_(j, i) = aij;
      for (int k = 1; k <= n; ++k) {
        const REAL zki = Z_(k, i), zkj = Z_(k, j);
        Z_(k, i) = c * zki - s * zkj;
        Z_(k, j) = s * zki + c * zkj;
      }
      stable = 0;
    }
    if (stable) return 0;
  }
  return 1;
}

/* Classical Jacobi of the reference's CPU prototype jac.cpp:60-217.  Column
 * l keeps the row maxpos[l] of its largest off-diagonal |a(k,l)| and that
 * value absele[l] (jac.cpp:82-92); each step takes the column j with the
 * largest absele and rotates the pair (i = maxpos[j], j) (find_maxpos,
 * jac.cpp:40-55).  The rotation is jac.cpp's: t from tau = beta/alpha or
 * alpha/beta by the larger of |alpha| = |2 a_ij| and |beta| = |a_jj - a_ii|
 * (jac.cpp:116-129), c = 1/sqrt(1+t^2), s = c t; w(i), w(j) and a(i,j) as
 * jac.cpp:143-146; rows/columns i, j of A and columns i, j of Z rotated
 * (jac.cpp:151-177); column maxima raised as jac.cpp:188-212.  Stops when the
 * pivot repeats (jac.cpp:105) or when absele[i] < (|a_ii| + |a_jj|) tol
 * (jac.cpp:114).  Two guards are added: when every tracked maximum is zero
 * (the matrix is diagonal: jac.cpp would then index row 0), and after
 * max_rot rotations (returns 1).  Eigenvalues unsorted in w. */
static int FN(jacobi_classical)(int nm, int n, REAL *a, REAL *z, REAL *w, REAL tol, long max_rot, int *maxpos,
                                REAL *absele)
{
  FN(zero_identity)(nm, n, z);
  for (int l = 1; l <= n; ++l) w[l - 1] = A_(l, l);
  for (int l = 1; l <= n; ++l) {
    int pos = 0;
    REAL amax = 0;
    for (int k = 1; k <= n; ++k)
      if (k != l && FN(fabs_)(A_(k, l)) > amax) { amax = FN(fabs_)(A_(k, l)); pos = k; }
    maxpos[l - 1] = pos;
    absele[l - 1] = amax;
  }
  int i = 0, j = 0;
  for (long itr = 0; itr < max_rot; ++itr) {
    const int i0 = i, j0 = j;
    int jm = 1;
    REAL b = 0;
    for (int k = 1; k <= n; ++k)
      if (b < absele[k - 1]) { b = absele[k - 1]; jm = k; }
    if (b == (REAL)0) return 0;
    i = maxpos[jm - 1];
    j = jm;
    if (i0 == i && j0 == j) return 0;
    REAL aij = A_(i, j);
    const REAL bij = absele[i - 1], aii = w[i - 1], ajj = w[j - 1];
    if (bij < (FN(fabs_)(aii) + FN(fabs_)(ajj)) * tol) return 0;
    const REAL alpha = 2 * aij, beta = ajj - aii;
    REAL t;
    if (FN(fabs_)(alpha) >= FN(fabs_)(beta)) {
      const REAL tau = beta / alpha, r = FN(sqrt_)(1 + tau * tau);
      t = 1 / (tau + (tau >= 0 ? r : -r));
    } else {
      const REAL tau = alpha / beta, r = FN(sqrt_)(1 + tau * tau);
      t = tau / (1 + r);
    }
    const REAL c = 1 / FN(sqrt_)(1 + t * t), sn = c * t;
    w[i - 1] = (c * c) * aii - 2 * (c * sn) * aij + (sn * sn) * ajj;
    w[j - 1] = (sn * sn) * aii + 2 * (c * sn) * aij + (c * c) * ajj;
    aij = (c * sn) * (aii - ajj) + (c - sn) * (c + sn) * aij;
    A_(i, j) = aij;
    A_(j, i) = aij;
    absele[i - 1] = absele[j - 1] = FN(fabs_)(aij);
    for (int k = 1; k <= n; ++k) {
      const REAL zki = Z_(k, i), zkj = Z_(k, j);
      Z_(k, i) = c * zki - sn * zkj;
      Z_(k, j) = sn * zki + c * zkj;
      if (k != i && k != j) {
        const REAL aki = A_(k, i), akj = A_(k, j);
        const REAL bik = c * aki - sn * akj, bjk = sn * aki + c * akj;
        A_(i, k) = A_(k, i) = bik;
        A_(j, k) = A_(k, j) = bjk;
        if (maxpos[k - 1] == i) absele[k - 1] = FN(fabs_)(bik);
        if (maxpos[k - 1] == j) absele[k - 1] = FN(fabs_)(bjk);
      }
    }
    for (int k = 1; k <= n; ++k) {
      if (k == i || k == j) continue;
      const REAL bik = FN(fabs_)(A_(i, k)), bjk = FN(fabs_)(A_(j, k));
      if (absele[i - 1] < bik) { maxpos[i - 1] = k; absele[i - 1] = bik; }
      if (absele[j - 1] < bjk) { maxpos[j - 1] = k; absele[j - 1] = bjk; }
      if (absele[k - 1] < bik) { maxpos[k - 1] = i; absele[k - 1] = bik; }
      if (absele[k - 1] < bjk) { maxpos[k - 1] = j; absele[k - 1] = bjk; }
    }
  }
  return 1;
}

/* Solver ids (shared with the GPU library's extended entry points; 3 is the
 * CPU-only classical Jacobi). */
#ifndef ORACLE_SOLVERS
#define ORACLE_SOLVERS
enum { ORC_IMTQL2 = 0, ORC_TQL2 = 1, ORC_JACOBI = 2, ORC_JACOBI_CLASSICAL = 3 };
#endif

/* One matrix, full pipeline as eigen_GPU_batch does (eigen_GPU_batch.cu:93-134):
 * tridiagonalise, QL with eigenvectors (z = work area), sort, back-transform,
 * then a(:,j) <- z(:,pos(j)) and w ascending.  Returns the QL error (0 = ok).
 * On error w holds the unsorted diagonal and a is left in tridiagonal storage,
 * as the reference does (the back-transform is skipped, eigen_GPU_batch.cu:128).
 * work: at least nm*n + 3n + n REAL plus n ints (allocated by the caller). */
static int FN(eig1)(int nm, int n, REAL *a, REAL *w, int solver, REAL *z, REAL *work, int *pos)
{
  REAL *d = work, *e = work + n, *tmp = work + 2 * n;
  const REAL eps = (sizeof(REAL) == 8) ? (REAL)DBL_EPSILON : (REAL)FLT_EPSILON;
  const REAL tol = eps * (REAL)((sizeof(REAL) == 8) ? 512 : 16);
  if (n == 1) { w[0] = a[0]; a[0] = 1; return 0; }
  if (solver == ORC_JACOBI_CLASSICAL) {
    /* jac.cpp reads the full symmetric matrix: mirror the upper triangle.
     * Tolerance eps^2 as jac.cpp's driver (jac.cpp:227-228). */
    for (int c = 1; c <= n; ++c) for (int r = c + 1; r <= n; ++r) A_(r, c) = A_(c, r);
    const int err = FN(jacobi_classical)(nm, n, a, z, d, eps * eps, 200L * n * n, pos, tmp);
    FN(sort_pos)(n, d, pos);
    for (int j = 0; j < n; ++j) w[j] = d[j];
    for (int c = 1; c <= n; ++c) for (int r = 1; r <= n; ++r) A_(r, c) = Z_(r, pos[c - 1] + 1);
    return err;
  }
  if (solver == ORC_JACOBI) {
    /* Jacobi reads the full symmetric matrix: mirror the upper triangle. */
    for (int c = 1; c <= n; ++c) for (int r = c + 1; r <= n; ++r) A_(r, c) = A_(c, r);
    const REAL jtol = (sizeof(REAL) == 8) ? (REAL)1e-20 : (REAL)1e-10;
    const int err = FN(jacobi)(nm, n, a, z, d, jtol, 10000);
    FN(sort_pos)(n, d, pos);
    for (int j = 0; j < n; ++j) w[j] = d[j];
    for (int c = 1; c <= n; ++c) for (int r = 1; r <= n; ++r) A_(r, c) = Z_(r, pos[c - 1] + 1);
    return err;
  }
  FN(tred1)(nm, n, a, d, e, tmp);
  const int ierr = (solver == ORC_TQL2) ? FN(tql2)(nm, n, d, e, z, tol, 100)
                                        : FN(imtql2)(nm, n, d, e, z, tol, 100);
  if (ierr) { for (int j = 0; j < n; ++j) w[j] = d[j]; return ierr; }
  FN(sort_pos)(n, d, pos);
  for (int j = 0; j < n; ++j) w[j] = d[j];
  FN(trbak)(nm, n, a, z);
  for (int c = 1; c <= n; ++c) for (int r = 1; r <= n; ++r) A_(r, c) = Z_(r, pos[c - 1] + 1);
  return 0;
}

/* Batched driver: matrix k at a + k*nm*n, eigenvalues at w + k*n, one matrix per
 * OpenMP thread (static schedule).  info[k] = QL error of matrix k (may be NULL).
 * Returns the number of matrices with a nonzero error, or -1 on bad arguments
 * (L<1, n<1, nm<n), mirroring eigen_GPU_batch.cu:252. */
int FN(oracle_eig_batch)(int L, int nm, int n, REAL *a, REAL *w, int solver, int *info, int nthreads)
{
  if (L < 1 || n < 1 || nm < n) return -1;
  int nbad = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : nbad)
  {
    REAL *z = (REAL *)malloc(sizeof(REAL) * ((size_t)nm * n + 4 * (size_t)n + 8));
    int *pos = (int *)malloc(sizeof(int) * (size_t)(n + 1));
#pragma omp for schedule(static)
    for (int k = 0; k < L; ++k) {
      REAL *ak = a + (size_t)k * nm * n;
      REAL *wk = w + (size_t)k * n;
      const int err = FN(eig1)(nm, n, ak, wk, solver, z, z + (size_t)nm * n, pos);
      if (info) info[k] = err;
      if (err) nbad++;
    }
    free(z);
    free(pos);
  }
  return nbad;
}

/* Tridiagonal stage alone (for tests of the storage convention). */
void FN(oracle_tred1)(int nm, int n, REAL *a, REAL *d, REAL *e)
{
  REAL *tmp = (REAL *)malloc(sizeof(REAL) * (size_t)(2 * n + 2));
  FN(tred1)(nm, n, a, d, e, tmp);
  free(tmp);
}

/* QL stage alone on a given tridiagonal (d, e with e(n)=0): z = eigenvectors of T
 * (unsorted), d = eigenvalues.  variant 0 = imtql2, 1 = tql2. */
int FN(oracle_ql)(int nm, int n, REAL *d, REAL *e, REAL *z, int variant)
{
  const REAL eps = (sizeof(REAL) == 8) ? (REAL)DBL_EPSILON : (REAL)FLT_EPSILON;
  const REAL tol = eps * (REAL)((sizeof(REAL) == 8) ? 512 : 16);
  return variant ? FN(tql2)(nm, n, d, e, z, tol, 100) : FN(imtql2)(nm, n, d, e, z, tol, 100);
}

/* Accuracy metrics of eigen_GPU_check.cu:34-81, in double, per matrix:
 *   out[0] = err1 = ||A Z - Z D||_F over the first m eigenpairs (full A read)
 *   out[1] = err2 = ||Z^T Z - I||_F over the first m vectors
 *   out[2] = TOL  = eps*(2*512-1 | 2*16-1)*sqrt(n+16)
 *   out[3] = pass flags: bit0 residual fails (err1 > max(|d1|,|dn|)*TOL),
 *                        bit1 orthogonality fails (err2 > n*TOL)
 *   out[4] = max_j ||A z_j - d_j z_j||_2 / ||A||_F  (per-eigenpair, extra)
 *   out[5] = err1 / ||A||_F */
void FN(oracle_check)(int nm, int n, int m, const REAL *a, const REAL *w, const REAL *z, double *out)
{
  double err1 = 0, err2 = 0, anrm = 0, worst = 0;
  for (int c = 1; c <= n; ++c) for (int r = 1; r <= n; ++r) { const double t = A_(r, c); anrm += t * t; }
  anrm = sqrt(anrm);
  for (int i = 1; i <= m; ++i) {
    double col = 0;
    for (int k = 1; k <= n; ++k) {
      double ek = 0;
      for (int j = 1; j <= n; ++j) ek += (double)A_(k, j) * (double)Z_(j, i);
      const double r = ek - (double)w[i - 1] * (double)Z_(k, i);
      col += r * r;
    }
    err1 += col;
    if (col > worst) worst = col;
  }
  err1 = sqrt(err1);
  for (int i = 1; i <= m; ++i)
    for (int j = i; j <= m; ++j) {
      double r = 0;
      for (int k = 1; k <= n; ++k) r += (double)Z_(k, i) * (double)Z_(k, j);
      r -= (i == j) ? 1.0 : 0.0;
      err2 += ((i == j) ? 1.0 : 2.0) * r * r;
    }
  err2 = sqrt(err2);
  const double epsT = (sizeof(REAL) == 8) ? 2.220446049250313e-16 : 1.1920928955078125e-07;
  const double TOL = epsT * ((sizeof(REAL) == 8) ? 1023.0 : 31.0) * sqrt((double)(n + 16));
  const double dmax = fmax(fabs((double)w[0]), fabs((double)w[n - 1]));
  int flags = 0;
  if (err1 > dmax * TOL) flags |= 1;
  if (err2 > n * TOL) flags |= 2;
  out[0] = err1;
  out[1] = err2;
  out[2] = TOL;
  out[3] = flags;
  out[4] = (anrm > 0) ? sqrt(worst) / anrm : sqrt(worst);
  out[5] = (anrm > 0) ? err1 / anrm : err1;
}

#undef FN
#undef A_
#undef Z_
