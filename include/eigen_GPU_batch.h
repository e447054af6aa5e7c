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
 *   m  : 1 <= m <= n.  All n eigenpairs are computed and returned for any m.
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
  EIGEN_GPU_BATCH_JACOBI = 2  /* two-sided Jacobi on the full matrix (jac_tiled.hpp) */
};

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

/* Library version string ("eigen-mi355x <major>.<minor>"). */
const char *eigen_GPU_batch_version(void);

#ifdef __cplusplus
}
#endif

#endif /* EIGEN_GPU_BATCH_H_MI355X */
