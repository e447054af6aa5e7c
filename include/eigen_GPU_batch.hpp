/*
 * eigen_GPU_batch.hpp - C++ interface of the MI355X batched symmetric eigensolver.
 *
 * Drop-in for the reference's header-only templates (eigen_GPU_batch.hpp:104-157):
 *   eigen_GPU_batch_BufferSize<T>(L, nm, n, m, a, w, &lwork)
 *   eigen_GPU_batch<T>(L, nm, n, m, a, w, wk, stream)
 * Unlike the reference, the buffer size comes from the library itself
 * (eigen_GPU_batch_bufsize), so the size the caller allocates and the size
 * the kernels use can never disagree (an autotuned choice is only run when it
 * fits the size this query reports), and no device query is made.
 */
#ifndef EIGEN_GPU_BATCH_HPP_MI355X
#define EIGEN_GPU_BATCH_HPP_MI355X

#include <type_traits>
#include "eigen_GPU_batch.h"

typedef hipError_t gpuError_t;

/* Replaces eigen_GPU_batch_get_Launch_params<T> (reference eigen_GPU_batch.hpp:23-102):
 * the launch shape of the library's first kernel for (L, n, T) with the default
 * solver (eigen_GPU_batch_launch_params); all outputs 0 and an error for invalid
 * arguments, as in the reference. */
template <class T>
inline gpuError_t eigen_GPU_batch_get_Launch_params(const int L, const int n, int &grpSZ_, int &numTB_,
                                                    int &numGR_, int &numTH_, int &sizeSH_, int &sizeL2_)
{
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                "eigen_GPU_batch supports float and double");
  int p[6] = {0, 0, 0, 0, 0, 0};
  const int rc = eigen_GPU_batch_launch_params(L, n, (int)sizeof(T), EIGEN_GPU_BATCH_IMTQL2, p);
  grpSZ_ = p[0]; numTB_ = p[1]; numGR_ = p[2]; numTH_ = p[3]; sizeSH_ = p[4]; sizeL2_ = p[5];
  return (gpuError_t)rc;
}

/* Replaces eigen_GPU_batch_BufferSize<T> (reference eigen_GPU_batch.hpp:104-118). */
template <class T>
inline gpuError_t eigen_GPU_batch_BufferSize(const int L, const int nm, const int n, const int m,
                                             T * /*a*/, T * /*w*/, size_t *lwork)
{
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                "eigen_GPU_batch supports float and double");
  *lwork = 0;
  if (L < 1 || nm < n || n < 1 || m < 1 || m > n) return hipErrorInvalidValue;
  *lwork = eigen_GPU_batch_bufsize(L, nm, n, m, (int)sizeof(T), EIGEN_GPU_BATCH_IMTQL2);
  return hipSuccess;
}

/* Replaces eigen_GPU_batch<T> (reference eigen_GPU_batch.hpp:120-157). */
template <class T>
inline void eigen_GPU_batch(const int L, const int nm, const int n, const int m,
                            T *a, T *w, T *wk, const gpuStream_t stream = NULL)
{
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                "eigen_GPU_batch supports float and double");
  if (std::is_same<T, float>::value)
    eigen_GPU_batch_FP(L, nm, n, m, (float *)a, (float *)w, (float *)wk, stream);
  else
    eigen_GPU_batch_DP(L, nm, n, m, (double *)a, (double *)w, (double *)wk, stream);
}

#endif /* EIGEN_GPU_BATCH_HPP_MI355X */
