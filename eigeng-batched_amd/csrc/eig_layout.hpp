// eig_layout.hpp - the interleaved batch layout (element e of matrix k at
// a[k + L e], the matrix index fastest; the reference's to-do "flexible
// interleave data format", README:166) converted to and from the batch-strided
// layout the solver kernels read (matrix k at a + k nm n).  Both directions are
// the transpose of an R x C row-major array: out[j R + i] = in[i C + j], done
// in 64 x 64 tiles through LDS so that reads and writes are both coalesced.
#pragma once
#include <hip/hip_runtime.h>

namespace eigmi {

template <typename T>
__global__ __launch_bounds__(256) void eig_transpose_kernel(long long R, long long C, const T *__restrict__ in,
                                                            T *__restrict__ out)
{
  __shared__ T tile[64][65];
  const long long tr = (R + 63) / 64, tc = (C + 63) / 64;
  const int x = threadIdx.x & 63, y = threadIdx.x >> 6;
  for (long long t = blockIdx.x; t < tr * tc; t += gridDim.x) {
    const long long i0 = (t % tr) * 64, j0 = (t / tr) * 64;
    // read rows i0.., columns j0.. (contiguous along j)
#pragma unroll 4
    for (int k = 0; k < 16; ++k) {
      const int ii = y + 4 * k;
      const long long i = i0 + ii, j = j0 + x;
      if (i < R && j < C) tile[ii][x] = in[i * C + j];
    }
    __syncthreads();
    // write out rows j0.. (contiguous along i)
#pragma unroll 4
    for (int k = 0; k < 16; ++k) {
      const int jj = y + 4 * k;
      const long long j = j0 + jj, i = i0 + x;
      if (i < R && j < C) out[j * R + i] = tile[x][jj];
    }
    __syncthreads();
  }
}

} // namespace eigmi
