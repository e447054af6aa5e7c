// eig_warmnet.hpp - the warm-start network's kernels: a 3-layer perceptron that
// maps the upper triangle of a symmetric matrix (n <= 64) to an approximate
// eigenbasis Q0 (n x n), trained on the plain solver's eigenpairs.
//
//   x  = features(A): the upper triangle (column by column) of (A - mu I) / s,
//        mu = trace(A) / n, s the RMS of the upper triangle of A - mu I
//        (the eigenvectors do not depend on mu or s)
//   h1 = relu(W1^T x + b1), h2 = relu(W2^T h1 + b2), y = W3^T h2 + b3,
//   Q0 = y as an n x n column-major matrix (re-orthonormalised by eig_qr_kernel)
//
// Loss: mean squared error against the target eigenvectors with each column's
// sign fixed so that its entry of largest magnitude is positive (eigenvectors
// are defined up to sign).  The products run on the matrix cores through
// eig_gemm_kernel (FP32 operands); these kernels are the element-wise and
// reduction steps around them and the Adam update.  Parameters are one FP32
// array: W1 (Din x H), b1, W2 (H x H), b2, W3 (H x n^2), b3, matrices
// column-major, Din = n (n + 1) / 2.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace eigmi {

struct WnDims {
  int n, din, h, dout;
  long long w1, b1, w2, b2, w3, b3, count;  // offsets into the parameter array
};

__host__ __device__ inline WnDims wn_dims(int n, int h)
{
  WnDims d;
  d.n = n;
  d.din = n * (n + 1) / 2;
  d.h = h;
  d.dout = n * n;
  d.w1 = 0;
  d.b1 = d.w1 + (long long)d.din * h;
  d.w2 = d.b1 + h;
  d.b2 = d.w2 + (long long)h * h;
  d.w3 = d.b2 + h;
  d.b3 = d.w3 + (long long)h * d.dout;
  d.count = d.b3 + d.dout;
  return d;
}

// splitmix64 of a counter, then Box-Muller: a standard normal per index
__device__ inline float wn_normal(unsigned long long seed, unsigned long long i)
{
  auto mix = [](unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  const unsigned long long a = mix(seed ^ mix(2 * i)), b = mix(seed ^ mix(2 * i + 1));
  const float u1 = ((a >> 40) + 0.5f) * (1.0f / 16777216.0f);
  const float u2 = (b >> 40) * (1.0f / 16777216.0f);
  return sqrtf(-2.0f * logf(u1)) * cosf(6.283185307179586f * u2);
}

// He-normal weights (std sqrt(2 / fan_in)), zero biases
__global__ __launch_bounds__(256) void wn_init_kernel(WnDims d, float *__restrict__ p, unsigned long long seed)
{
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < d.count; i += (long long)gridDim.x * 256) {
    float v = 0.0f;
    if (i < d.b1) v = wn_normal(seed, i) * sqrtf(2.0f / d.din);
    else if (i >= d.w2 && i < d.b2) v = wn_normal(seed, i) * sqrtf(2.0f / d.h);
    else if (i >= d.w3 && i < d.b3) v = wn_normal(seed, i) * sqrtf(2.0f / d.h);
    p[i] = v;
  }
}

// features of matrix b (upper triangle of a, ld nm) into x[:, b] (ld din)
__global__ __launch_bounds__(256) void wn_features_kernel(int L, int nm, int n, const float *__restrict__ a,
                                                          float *__restrict__ x)
{
  __shared__ float red[2][4];
  const int t = threadIdx.x, din = n * (n + 1) / 2;
  for (long long b = blockIdx.x; b < L; b += gridDim.x) {
    const float *__restrict__ A = a + (size_t)b * nm * n;
    float tr = 0.0f;
    for (int i = t; i < n; i += 256) tr += A[i + (size_t)i * nm];
    for (int o = 32; o >= 1; o >>= 1) tr += __shfl_xor(tr, o, 64);
    if ((t & 63) == 0) red[0][t >> 6] = tr;
    __syncthreads();
    const float mu = (red[0][0] + red[0][1] + red[0][2] + red[0][3]) / n;
    float ss = 0.0f;
    for (int e = t; e < din; e += 256) {
      int j = (int)((sqrtf(8.0f * e + 1.0f) - 1.0f) * 0.5f);
      while (j * (j + 1) / 2 > e) --j;
      while ((j + 1) * (j + 2) / 2 <= e) ++j;
      const int i = e - j * (j + 1) / 2;
      const float v = A[i + (size_t)j * nm] - (i == j ? mu : 0.0f);
      ss += v * v;
    }
    for (int o = 32; o >= 1; o >>= 1) ss += __shfl_xor(ss, o, 64);
    if ((t & 63) == 0) red[1][t >> 6] = ss;
    __syncthreads();
    const float rms = sqrtf((red[1][0] + red[1][1] + red[1][2] + red[1][3]) / din);
    const float inv = rms > 0.0f ? 1.0f / rms : 1.0f;
    float *__restrict__ xb = x + (size_t)b * din;
    for (int e = t; e < din; e += 256) {
      int j = (int)((sqrtf(8.0f * e + 1.0f) - 1.0f) * 0.5f);
      while (j * (j + 1) / 2 > e) --j;
      while ((j + 1) * (j + 2) / 2 <= e) ++j;
      const int i = e - j * (j + 1) / 2;
      xb[e] = (A[i + (size_t)j * nm] - (i == j ? mu : 0.0f)) * inv;
    }
    __syncthreads();
  }
}

// z[r + c ld] += bias[r], then relu if asked (rows x cols)
__global__ __launch_bounds__(256) void wn_bias_act_kernel(int rows, long long cols, float *__restrict__ z, long long ld,
                                                          const float *__restrict__ bias, int relu)
{
  const long long total = (long long)rows * cols;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const long long c = e / rows;
    const int r = (int)(e - c * rows);
    float v = z[r + c * ld] + bias[r];
    if (relu) v = v > 0.0f ? v : 0.0f;
    z[r + c * ld] = v;
  }
}

// dz *= (act > 0)
__global__ __launch_bounds__(256) void wn_relu_grad_kernel(long long count, float *__restrict__ dz,
                                                           const float *__restrict__ act)
{
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < count; e += (long long)gridDim.x * 256)
    if (!(act[e] > 0.0f)) dz[e] = 0.0f;
}

// y (n^2 x B, ld n^2) := dL/dy of the mean squared error against the
// sign-fixed targets v (matrix b at v + b nm n, column j rows 0..n-1); the
// loss (sum of squares / (B n^2)) is added to *loss.  One thread per column.
__global__ __launch_bounds__(256) void wn_mse_grad_kernel(int B, int nm, int n, float *__restrict__ y,
                                                          const float *__restrict__ v, float *__restrict__ loss)
{
  __shared__ float red[4];
  const long long cols = (long long)B * n;
  const float scale = 1.0f / ((float)B * n * n);
  float part = 0.0f;
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < cols; c += (long long)gridDim.x * 256) {
    const long long b = c / n;
    const int j = (int)(c - b * n);
    const float *__restrict__ tv = v + (size_t)b * nm * n + (size_t)j * nm;
    float best = 0.0f;
    for (int i = 0; i < n; ++i) {
      const float x = tv[i];
      if (fabsf(x) > fabsf(best)) best = x;
    }
    const float sg = best < 0.0f ? -1.0f : 1.0f;
    float *__restrict__ yc = y + (size_t)b * n * n + (size_t)j * n;
    for (int i = 0; i < n; ++i) {
      const float d = yc[i] - sg * tv[i];
      part += d * d;
      yc[i] = 2.0f * d * scale;
    }
  }
  for (int o = 32; o >= 1; o >>= 1) part += __shfl_xor(part, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (red[0] + red[1] + red[2] + red[3]) * scale);
}

// db[r] += sum over the columns c of z[r + c ld] (workgroups split the columns;
// db must be cleared first)
__global__ __launch_bounds__(256) void wn_rowsum_kernel(int rows, long long cols, const float *__restrict__ z,
                                                        long long ld, float *__restrict__ db, int csplit)
{
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= rows) return;
  const long long per = (cols + csplit - 1) / csplit;
  const long long c0 = blockIdx.y * per, c1 = c0 + per < cols ? c0 + per : cols;
  float s = 0.0f;
  for (long long c = c0; c < c1; ++c) s += z[r + c * ld];
  atomicAdd(db + r, s);
}

__global__ __launch_bounds__(256) void wn_fill_kernel(long long count, float *__restrict__ p, float v)
{
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < count; e += (long long)gridDim.x * 256) p[e] = v;
}

// Adam (Kingma & Ba 2015) with bias correction, step >= 1
__global__ __launch_bounds__(256) void wn_adam_kernel(long long count, float *__restrict__ p,
                                                      const float *__restrict__ g, float *__restrict__ m,
                                                      float *__restrict__ v, float lr, float b1, float b2, float eps,
                                                      float c1, float c2)
{
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < count; e += (long long)gridDim.x * 256) {
    const float ge = g[e];
    const float me = b1 * m[e] + (1.0f - b1) * ge;
    const float ve = b2 * v[e] + (1.0f - b2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    p[e] -= lr * (me * c1) / (sqrtf(ve * c2) + eps);
  }
}

// y (n^2 x B, ld n^2, matrix b column-major ld n) -> q (ld nm, matrix b at q + b nm n)
__global__ __launch_bounds__(256) void wn_store_q_kernel(int B, int nm, int n, const float *__restrict__ y,
                                                         float *__restrict__ q)
{
  const long long total = (long long)B * n * n;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const long long b = e / ((long long)n * n);
    const int r = (int)(e - b * n * n);
    const int j = r / n, i = r - j * n;
    q[(size_t)b * nm * n + (size_t)j * nm + i] = y[e];
  }
}

} // namespace eigmi
