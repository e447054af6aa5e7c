/*
 * eigen_GPU_batch.h - C ABI of the MI355X (gfx950) batched symmetric eigensolver.
 *
 * Drop-in for RIKEN-RCCS/EigenG-Batched's C interface (eigen_GPU_batch.h:26-30):
 * same symbols, argument order, types and batch-strided column-major layout.
 *
 *   a  : L matrices, matrix k at a + k*nm*n, element (r,c) at a[r + c*nm]
 *        (column-major, leading dimension nm >= n).  Only the upper triangle
 *        and the diagonal are read.  On return column j (rows 0..n-1) holds
 *        the unit eigenvector of w[k*n + j].
 *   w  : L*n eigenvalues, ascending within each matrix, matrix k at w + k*n.
 *   wk : device workspace of at least lwork bytes (see the query below).
 *   m  : 1 <= m <= n.  All n eigenvalues are returned in w; the eigenvectors
 *        of the first m of them are written to columns 0..m-1 of a, and columns
 *        m..n-1 of a are left unspecified (the reference passes m through but
 *        stores inconsistently for m < n: w offset pos*m, eigen_GPU_batch.cu:89,
 *        against stride n, :133; back-transform of m columns, hhtr2sy.hpp:116,144).
 *
 * Workspace query (eigen_GPU_batch.cu:329-359): calling with wk == NULL does not
 * compute; it writes the required workspace size in bytes as a size_t into the
 * host memory at w.  Unlike the reference, the size is nonzero for every valid
 * n (n == 1 included) and always a full size_t, so `query -> hipMalloc ->
 * compute` works for n == 1 as well.
 *
 * Invalid arguments (L<1, n<1, nm<n, m<1, m>n) make the call a no-op, as in the
 * reference (eigen_GPU_batch.cu:252).  The call is asynchronous on `stream`.
 * All pointers are caller-owned device memory; nothing is allocated here.
 */
#ifndef EIGEN_GPU_BATCH_H_MI355X
#define EIGEN_GPU_BATCH_H_MI355X

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The reference's stream typedef (gpu_arch.h:54); hipStream_t on this target. */
typedef hipStream_t gpuStream_t;

/* Replaces eigen_GPU_batch_DP (reference eigen_GPU_batch.h:27, eigen_GPU_batch.cu:329-344). */
void eigen_GPU_batch_DP(const int L, const int nm, const int n, const int m,
                        double *a, double *w, double *wk, const gpuStream_t stream);

/* Replaces eigen_GPU_batch_FP (reference eigen_GPU_batch.h:30, eigen_GPU_batch.cu:346-359). */
void eigen_GPU_batch_FP(const int L, const int nm, const int n, const int m,
                        float *a, float *w, float *wk, const gpuStream_t stream);

/*
 * Extensions (no reference symbol; the reference selects these at compile time
 * with USE_IMTQL / DO_SORT, eigen_GPU_batch.cu:1-11, and reports nothing).
 */

/* Solver selection for the _x entry points. */
enum eigen_GPU_batch_solver {
  EIGEN_GPU_BATCH_IMTQL2 = 0, /* Householder + implicit-shift QL + back-transform (default) */
  EIGEN_GPU_BATCH_TQL2 = 1,   /* Householder + explicit-shift QL (the reference's shipped QL) */
  EIGEN_GPU_BATCH_JACOBI = 2, /* two-sided Jacobi on the full matrix (jac_tiled.hpp) */
  /* FP64 entry points only, n <= 256: mixed precision.  The matrices rounded to
   * FP32 are solved by the FP32 path, then the FP32 eigenpairs are refined in
   * FP64 on the matrix cores (Ogita-Aishima iteration; clusters of up to 8
   * eigenvalues closer than the iteration's separation bound are resolved by
   * Rayleigh-Ritz); a matrix the refinement does not accept (larger cluster,
   * measures above 1e-6 before the last update, failed FP32 solve) is solved
   * by the plain FP64 workgroup solver.  The reference lists "Half precision
   * version + Iterative refinment" as future work (README:169). */
  EIGEN_GPU_BATCH_MIXED = 3
};
/* bits 28-29 of `solver` with EIGEN_GPU_BATCH_MIXED: refinement iterations
 * (1..3; 0 selects the default, 3). */
#define EIGEN_GPU_BATCH_MIXED_ITERS_SHIFT 28
/* with EIGEN_GPU_BATCH_MIXED and 64 < n <= 256: the FP32 stage takes the
 * large-order split path and back-transforms with blocked reflectors (compact
 * WY, blocks of 64: V <- V + U (T (U^T V))) on the FP16 matrix cores with FP32
 * accumulation (error-compensated split operands), the QL rotations replayed
 * from the identity; the FP64 refinement then wins back FP64 accuracy (the
 * reference's blocked back-transformation reduces with shuffles,
 * hhtr2sy.hpp:144-202). */
#define EIGEN_GPU_BATCH_MIXED_WY16 0x8000000

/*
 * Tuning bits OR-ed into `solver` of the _x entry points and of
 * eigen_GPU_batch_bufsize (the default picks by n, precision and L: for n <= 64
 * the split path from 2048 matrices, from one matrix for FP64 16 <= n <= 32 and
 * FP32 28 <= n <= 32, from 1024 for FP32 n >= 56, the fused kernel below):
 *   0x100  one fused kernel per matrix group;
 *   0x200  split n <= 64 path: register tridiagonalisation, one-lane-per-matrix
 *          QL recurrence writing a rotation log, rotation apply kernel;
 *   0x400  (with 0x200) shrink the rotation log so every matrix takes the
 *          split path's overflow fallback (a host-forced test of that path);
 *   0x1000 / 0x2000 / 0x4000  run only the split path's tridiagonalisation /
 *          QL / apply phase (L <= 16384): successive calls on one stream with
 *          one workspace compose the full solve, so each phase can be timed;
 *   0x8000 FP32 64 < n <= 128: the register split path (two-wave
 *          tridiagonalisation, lane-per-matrix QL, 64-row-block apply), the
 *          default there from 2048 matrices; 0x200 selects the large-order
 *          split path instead;
 *   bits 16-17 / 20-23  workgroup size / resident workgroup count of the
 *          workgroup kernels; bits 18-19 (1..3) divide the split paths' batch
 *          chunk (16384 matrices for n <= 128, 8192 beyond; 2048 for the
 *          mixed solver's refinement) by 2, 4 or 8: the workspace shrinks by
 *          that factor and the latency-bound QL recurrence runs once more per
 *          chunk; bits 24-25 tridiagonalisation shape of the
 *          n > 64 split path; 0x4000000 (n <= 256) its full-storage form
 *          instead of the default lower-triangle form;
 *   0x40000000 eigenpairs in the order the QL iteration (or Jacobi) leaves
 *          them instead of ascending: the reference's DO_SORT 0
 *          (eigen_GPU_batch.cu:5,10; its README:57-58 calls unsorted the
 *          default).  Columns 0..m-1 then hold the first m of that order.
 */
#define EIGEN_GPU_BATCH_FUSED 0x100
#define EIGEN_GPU_BATCH_SPLIT 0x200
#define EIGEN_GPU_BATCH_SPLIT_TINYLOG 0x400
#define EIGEN_GPU_BATCH_PHASE_TRI 0x1000
#define EIGEN_GPU_BATCH_PHASE_QL 0x2000
#define EIGEN_GPU_BATCH_PHASE_APPLY 0x4000
#define EIGEN_GPU_BATCH_REG128 0x8000
#define EIGEN_GPU_BATCH_TRI_FULL 0x4000000
#define EIGEN_GPU_BATCH_UNSORTED 0x40000000
#define EIGEN_GPU_BATCH_CHUNK_SHIFT 18

/* Workspace size in bytes for (L, nm, n, m, solver); 0 for invalid arguments. */
size_t eigen_GPU_batch_bufsize(int L, int nm, int n, int m, int dtype_bytes, int solver);

/*
 * Same computation as eigen_GPU_batch_FP/_DP with an explicit solver and an
 * optional per-matrix status array `info` (device int[L], may be NULL):
 * info[k] = 0 on success, l >= 1 if the QL iteration for eigenvalue l did not
 * converge in 100 sweeps (eigenvalues then unsorted, eigenvectors not written),
 * as the reference's ierror (tql2_tiled.hpp:138).  Returns hipSuccess, or
 * hipErrorInvalidValue for invalid arguments (nothing launched), or the launch
 * error.  wk must hold eigen_GPU_batch_bufsize(...) bytes.
 */
int eigen_GPU_batch_FP_x(int L, int nm, int n, int m, float *a, float *w, void *wk,
                         hipStream_t stream, int solver, int *info);
int eigen_GPU_batch_DP_x(int L, int nm, int n, int m, double *a, double *w, void *wk,
                         hipStream_t stream, int solver, int *info);

/*
 * Uneven batches (no reference symbol: the reference solves one order per call,
 * eigen_GPU_batch.cu:248-325, so a mixed batch costs one call per order).
 *
 *   n    : device int[L], order of matrix k (1 <= n[k] <= nmax).
 *   nmax : host-side upper bound of the orders, 1 <= nmax <= 512.
 *   aoff, woff : device int64[L] element offsets of matrix k in a (column-major,
 *          leading dimension n[k]) and of its n[k] eigenvalues in w; both NULL
 *          means the packed layout (matrix k follows matrix k-1, offsets are the
 *          prefix sums of n[j]^2 and n[j]).
 *   info : device int[L] or NULL; as for the _x entry points, and -3 for a
 *          matrix whose order is outside 1..nmax (left untouched).
 *   wk   : device workspace of eigen_GPU_batch_uneven_bufsize(L, nmax, ...)
 *          bytes.
 * Solvers: IMTQL2 and TQL2.  Every step is a kernel launched on `stream`
 * (no allocation, memset or host synchronisation), so a call can be captured
 * in a HIP graph.  Returns hipSuccess, hipErrorInvalidValue or
 * hipErrorNotSupported (nothing launched), or the launch error.
 */
size_t eigen_GPU_batch_uneven_bufsize(int L, int nmax, int dtype_bytes, int solver);
int eigen_GPU_batch_uneven_FP(int L, int nmax, const int *n, const long long *aoff,
                              const long long *woff, float *a, float *w, void *wk,
                              hipStream_t stream, int solver, int *info);
int eigen_GPU_batch_uneven_DP(int L, int nmax, const int *n, const long long *aoff,
                              const long long *woff, double *a, double *w, void *wk,
                              hipStream_t stream, int solver, int *info);

/*
 * Accuracy check of a batch of eigenpairs, in double, on the GPU.
 * a: the full symmetric input matrices (leading dimension nm), w: eigenvalues,
 * z: eigenvectors (leading dimension nm), m: number of eigenpairs checked.
 * err1 = ||A Z - Z D||_F, err2 = ||Z^T Z - I||_F, TOL = eps (2*512-1 | 2*16-1)
 * sqrt(n+16); a matrix fails if err1 > max(|w_1|, |w_n|) TOL or err2 > n TOL.
 */
/* Replaces eigen_GPU_check_DP / _FP (reference eigen_GPU_check.h:20-23,
 * eigen_GPU_check.cu:186-198): the report (err1/max|w| or 0, err2 or 0, TOL,
 * eps, as floats) is written over the first four floats of each a_k, then read
 * back (the call synchronizes `stream`) and every failure printed to stdout.
 * When a matrix's storage (nm * n elements) is smaller than the 16-byte report
 * (FP32 n = 1 with nm <= 3, FP64 nm = n = 1) nothing is checked and a message
 * goes to stderr; the reference writes past the matrix there. */
void eigen_GPU_check_DP(const int L, const int nm, const int n, const int m,
                        double *a_, double *w_, double *z_, const gpuStream_t stream);
void eigen_GPU_check_FP(const int L, const int nm, const int n, const int m,
                        float *a_, float *w_, float *z_, const gpuStream_t stream);
/* Extension: asynchronous, inputs untouched; out (device double[4 L]) gets
 * (err1, err2, TOL, fail bits: 1 residual, 2 orthogonality) per matrix.
 * Returns hipSuccess, hipErrorInvalidValue (n > 512 or bad sizes) or the
 * launch error. */
int eigen_GPU_check_x_DP(int L, int nm, int n, int m, const double *a, const double *w,
                         const double *z, double *out, hipStream_t stream);
int eigen_GPU_check_x_FP(int L, int nm, int n, int m, const float *a, const float *w,
                         const float *z, double *out, hipStream_t stream);

/*
 * Autotuner (no reference symbol; the reference picks launch shapes by fixed
 * heuristics, eigen_GPU_batch.hpp:23-102).  Times every candidate tuning-bit
 * set for (L, nm, n, element type, solver) on the caller's batch: a_src holds
 * the L input matrices (left untouched), a is a work copy that is restored from
 * a_src before each trial and ends holding the eigenvectors of the chosen set,
 * w the eigenvalues.  The winner is returned in *best (its time in *best_ms;
 * pass it as tuning bits to the _x entry points, with a workspace sized by
 * eigen_GPU_batch_bufsize for those bits).  For this device the library
 * remembers the fastest candidate whose workspace fits in what the default
 * size query reports for the shape: later calls with the same n, element type
 * and solver, an L in the same power-of-two range and no tuning bits of their
 * own run it.  Size queries never change through tuning, so a workspace sized
 * before tuning stays large enough (the reference ABI carries no workspace
 * length).  cand / cand_ms (may be NULL) receive up to max_report candidates
 * and their times, *ncand their number.  wk must hold
 * eigen_GPU_batch_tune_bufsize(...) bytes.  Synchronizes stream.
 */
size_t eigen_GPU_batch_tune_bufsize(int L, int nm, int n, int dtype_bytes, int solver);
int eigen_GPU_batch_tune_FP(int L, int nm, int n, const float *a_src, float *a, float *w,
                            void *wk, hipStream_t stream, int solver, int *best, float *best_ms,
                            int max_report, int *cand, float *cand_ms, int *ncand);
int eigen_GPU_batch_tune_DP(int L, int nm, int n, const double *a_src, double *a, double *w,
                            void *wk, hipStream_t stream, int solver, int *best, float *best_ms,
                            int max_report, int *cand, float *cand_ms, int *ncand);
/* Remember tuning bits for (n, element size, solver, the power-of-two range of
 * L) on the current device, e.g. a choice saved from an earlier tuning run;
 * bits == 0 forgets the entry.  The same rule as for the tuner's own choices
 * holds: later calls without bits run them only when their workspace fits in
 * what the default size query reports.  Returns hipSuccess or
 * hipErrorInvalidValue (bad sizes, solver bits in `bits`, phase or test bits). */
int eigen_GPU_batch_tune_set(int L, int n, int dtype_bytes, int solver, int bits);
/* Forget every tuned choice (all devices). */
void eigen_GPU_batch_tune_reset(void);

/*
 * Products on either side of the solver (MFMA, FP32 / FP64 accumulation).
 * All are asynchronous on stream and return hipSuccess, hipErrorInvalidValue
 * (L < 0, n < 1, n > 512, nm < n, a null pointer, overlapping input and output)
 * or the launch error; L == 0 is a no-op.  Outputs are column-major with
 * leading dimension nm, matrix k at c + k nm n; rows n..nm-1 are not written;
 * both triangles are written and are exactly symmetric.
 *
 * symdmt: C_k = Z_k diag(d_k) Z_k^T, Z_k at z + k nm n, d_k at d + k n.
 *   Replaces the reference's single-matrix symdmt_batch experiment
 *   (symdmt.cu:5-87), batched.
 * matpow: C_k = scale Z_k diag(w_k)^p Z_k^T, i.e. a function of the matrix
 *   from the solver's eigenpairs (README:171 "any power of the input matrix"):
 *   p = -1 gives the inverse, p = 0.5 / -0.5 the square root and its inverse
 *   (an eigenvalue <= 0 with a non-integer p gives NaN, as pow does).
 */
int eigen_GPU_batch_symdmt_FP(int L, int nm, int n, const float *z, const float *d, float *c,
                              hipStream_t stream);
int eigen_GPU_batch_symdmt_DP(int L, int nm, int n, const double *z, const double *d, double *c,
                              hipStream_t stream);
int eigen_GPU_batch_matpow_FP(int L, int nm, int n, const float *z, const float *w, float p,
                              float scale, float *c, hipStream_t stream);
int eigen_GPU_batch_matpow_DP(int L, int nm, int n, const double *z, const double *w, double p,
                              double scale, double *c, hipStream_t stream);
/*
 * gram: A_k = alpha I + Y_k^T diag(r_k) Y_k, the LETKF matrix
 *   ((k-1)/rho) I + Y^T R^-1 Y formed on the GPU (the reference forms its test
 *   matrices on the host, main.cpp:63-195; no reference symbol).  Y_k: p x n,
 *   column-major, leading dimension ldy >= p, at y + k ldy n.  r_k = rinv +
 *   k rstride (rinv NULL: R = I; rstride 0: one vector for the whole batch).
 */
int eigen_GPU_batch_gram_FP(int L, int nm, int n, int p, const float *y, int ldy,
                            const float *rinv, int rstride, float alpha, float *a,
                            hipStream_t stream);
int eigen_GPU_batch_gram_DP(int L, int nm, int n, int p, const double *y, int ldy,
                            const double *rinv, int rstride, double alpha, double *a,
                            hipStream_t stream);

/*
 * Launch shape of the first kernel of the path a call with (L, n, element
 * size, solver) takes (tuning included), in the terms of the reference's
 * eigen_GPU_batch_get_Launch_params (eigen_GPU_batch.hpp:23-102): params[0]
 * lanes per matrix group, [1] workgroups, [2] matrices (groups) per
 * workgroup, [3] threads per workgroup, [4] dynamic LDS bytes passed at
 * launch (0: the kernels size their LDS statically), [5] L2 bytes of the
 * current device.  Returns hipSuccess, or hipErrorInvalidValue /
 * hipErrorNotSupported with params zeroed.
 */
int eigen_GPU_batch_launch_params(int L, int n, int dtype_bytes, int solver, int *params);

/*
 * No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The computation of eigen_GPU_batch_FP_x /
 * _DP_x (same kernels) with the workspace taken from a fixed device array owned
 * by the library: eigen_GPU_batch_noalloc_scratch_bytes() = 256 MiB on each
 * device that runs it, held until the process ends.  A batch whose workspace
 * exceeds the array is solved in consecutive sub-batches (halving the count
 * until it fits; a sub-batch may take another path than the whole batch would,
 * so results can then differ from the _x call's at rounding level).  Returns
 * hipErrorInvalidValue for bad sizes, phase or test bits, or when one matrix's
 * workspace exceeds the array; hipErrorNotSupported for n > 512.  Nothing is
 * allocated and no memset is issued; the array needs no clearing between
 * calls.  Calls that overlap on one device race on the array.
 */
size_t eigen_GPU_batch_noalloc_scratch_bytes(void);
int eigen_GPU_batch_FP_noalloc(int L, int nm, int n, int m, float *a, float *w, hipStream_t stream,
                               int solver, int *info);
int eigen_GPU_batch_DP_noalloc(int L, int nm, int n, int m, double *a, double *w, hipStream_t stream,
                               int solver, int *info);

/*
 * FP64 refinement of approximate eigenpairs (no reference symbol; the second
 * half of EIGEN_GPU_BATCH_MIXED, usable on eigenvectors from any source):
 * a holds the L matrices (upper triangle and diagonal read, ld nm), x their
 * approximate eigenvectors (n columns, ld ldx >= n, matrix k at x + k ldx n),
 * refined in place by `iters` (1..8) Ogita-Aishima iterations on the matrix
 * cores with Rayleigh-Ritz on clusters of up to 8 close eigenvalues; w gets
 * the eigenvalues, ascending, with the columns of x in the same order.
 * info[k] = 0 when the last update's first-order entries are at most 1e-6
 * (the result is then at FP64 rounding level), else a positive bit set:
 * 2 cluster beyond 8, 4 not converged, 8 non-finite values; x and w of such
 * a matrix are left as the iteration reached them.  n <= 256; workspace of
 * eigen_GPU_batch_refine_bufsize(L, n) bytes.  Returns hipSuccess,
 * hipErrorInvalidValue, hipErrorNotSupported or the launch error.
 */
size_t eigen_GPU_batch_refine_bufsize(int L, int n);
int eigen_GPU_batch_refine_DP(int L, int nm, int n, const double *a, double *x, int ldx, double *w, void *wk,
                              hipStream_t stream, int iters, int *info);

/*
 * Warm-started solve, n <= 64 (no reference symbol): q0 holds an approximate
 * eigenbasis per matrix (a's layout), e.g. the warm-start network's output.
 * Q = qr(q0) (Householder, R with a non-negative diagonal), B = Q^T A Q and
 * V = Q Z_B on the matrix cores (FP32 / FP64 operands), Z_B from cyclic
 * Jacobi on B (the reference's jac_tiled, jac_tiled.hpp:3-152, started from
 * Q instead of the identity) with at most max_sweeps (1..40) sweeps.  A
 * matrix whose Jacobi has not converged within max_sweeps is solved by the
 * plain solver (tridiagonalisation + QL) from its input.  Results as for the
 * _x entry points (sorted, first m eigenvector columns); sweeps (device int[L]
 * or NULL) gets the sweeps each Jacobi took (max_sweeps: fell back).  Workspace:
 * eigen_GPU_batch_warm_bufsize(L, n, dtype_bytes).
 */
size_t eigen_GPU_batch_warm_bufsize(int L, int n, int dtype_bytes);
int eigen_GPU_batch_warm_FP(int L, int nm, int n, int m, float *a, const float *q0, float *w, void *wk,
                            hipStream_t stream, int max_sweeps, int *info, int *sweeps);
int eigen_GPU_batch_warm_DP(int L, int nm, int n, int m, double *a, const double *q0, double *w, void *wk,
                            hipStream_t stream, int max_sweeps, int *info, int *sweeps);
/*
 * Warm-start network (no reference symbol): a 3-layer perceptron, hidden
 * width `hidden` (ReLU), from the upper triangle of (A - mu I) / s (mu =
 * trace / n, s the RMS of that triangle) to an n x n approximate eigenbasis
 * Q0 for eigen_GPU_batch_warm_FP, n <= 64.  Parameters are one FP32 array of
 * eigen_GPU_warmnet_params(n, hidden) floats: W1 (Din x H), b1, W2 (H x H), b2,
 * W3 (H x n^2), b3 (column-major, Din = n(n+1)/2).  The layers' products run
 * on the matrix cores (FP32 operands); wk holds the activations of `batch`
 * matrices (eigen_GPU_warmnet_bufsize(batch, n, hidden) bytes).
 *   init:    He-normal weights from `seed`, zero biases.
 *   forward: q0 (a's layout, ld nm) for L matrices, processed `batch` at a time.
 *   grad:    for B matrices a with target eigenvectors v (a's layout, columns
 *            ascending), the mean squared error of the output against v with
 *            each column's sign fixed (largest entry positive) into *loss
 *            (device float) and its gradient into grad (the parameters' layout).
 *   adam:    one Adam step (beta 0.9 / 0.999, eps 1e-8, bias-corrected; step >= 1).
 * All asynchronous on stream; return hipSuccess, hipErrorInvalidValue or the
 * launch error.
 */
size_t eigen_GPU_warmnet_params(int n, int hidden);
size_t eigen_GPU_warmnet_bufsize(int batch, int n, int hidden);
int eigen_GPU_warmnet_init(int n, int hidden, float *params, unsigned long long seed, hipStream_t stream);
int eigen_GPU_warmnet_forward(int L, int nm, int n, int hidden, const float *params, const float *a, float *q0,
                              int batch, void *wk, hipStream_t stream);
int eigen_GPU_warmnet_grad(int B, int nm, int n, int hidden, const float *params, const float *a, const float *v,
                           float *grad, float *loss, void *wk, hipStream_t stream);
int eigen_GPU_warmnet_adam(long long count, float *params, const float *grad, float *m, float *v, int step, float lr,
                           hipStream_t stream);

/* Q factor of the Householder QR of each of L matrices (n <= 64, ld nm, R with
 * a non-negative diagonal); q may equal a.  L == 0 is a no-op. */
int eigen_GPU_batch_qr_FP(int L, int nm, int n, const float *a, float *q, hipStream_t stream);
int eigen_GPU_batch_qr_DP(int L, int nm, int n, const double *a, double *q, hipStream_t stream);

/*
 * Interleaved batch layout (the reference's to-do "flexible interleave data
 * format", README:166; no reference symbol): element (r, c) of matrix k at
 * a[k + L (r + c nm)] and eigenvalue j of matrix k at w[k + L j], i.e. the
 * matrix index runs fastest.  Same arguments, options and results as the _x
 * entry points; the batch is transposed into the batch-strided layout inside
 * the workspace (eigen_GPU_batch_interleaved_bufsize bytes: the strided copy,
 * its eigenvalues and the _x workspace), solved, and transposed back.
 */
size_t eigen_GPU_batch_interleaved_bufsize(int L, int nm, int n, int m, int dtype_bytes, int solver);
int eigen_GPU_batch_interleaved_FP(int L, int nm, int n, int m, float *a, float *w, void *wk,
                                   hipStream_t stream, int solver, int *info);
int eigen_GPU_batch_interleaved_DP(int L, int nm, int n, int m, double *a, double *w, void *wk,
                                   hipStream_t stream, int solver, int *info);

/*
 * Batched products on the matrix cores (no reference symbol; the reference
 * reduces its back-transformation products with shuffles, hhtr2sy.hpp:144-202):
 *   C_b = alpha op(A_b) op(B_b) + beta D_b,  b = 0 .. batch-1,
 * column-major, op(X) = X (trans 0) or X^T (trans 1), op(A_b) M x K, op(B_b)
 * K x N, leading dimensions and batch strides in elements.  kind: 0 FP64
 * operands (v_mfma_f64_16x16x4f64), 1 FP32 (v_mfma_f32_16x16x4f32), 2 FP32
 * storage multiplied as FP16 (v_mfma_f32_16x16x16f16), 3 FP32 storage
 * multiplied as bfloat16 (v_mfma_f32_16x16x16bf16_1k), 4 FP32 storage split
 * into two halves (hi + lo / 2^11) and multiplied as three FP16 products with
 * FP32 accumulation, close to FP32 accuracy (error-compensated splitting,
 * Ootomo & Yokota 2022; |values| < 65504); FP32 accumulation for kinds 1-4.  D may be NULL (beta ignored) or equal to C (same ldc); C must not
 * overlap A or B.  Returns hipSuccess (batch == 0: nothing launched),
 * hipErrorInvalidValue, or the launch error.
 */
int eigen_GPU_batch_gemm(int kind, int transa, int transb, int M, int N, int K, double alpha,
                         const void *A, long long lda, long long stride_a,
                         const void *B, long long ldb, long long stride_b, double beta,
                         const void *D, long long ldd, long long stride_d,
                         void *C, long long ldc, long long stride_c, int batch, hipStream_t stream);

/* Library version string ("eigen-mi355x <major>.<minor>"). */
const char *eigen_GPU_batch_version(void);

#ifdef __cplusplus
}
#endif

#endif /* EIGEN_GPU_BATCH_H_MI355X */
