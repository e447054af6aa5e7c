// eig_gemm.hpp - strided-batched dense products on the matrix cores, for the
// steps of the solver that are real contractions:
//
//   C_b = alpha op(A_b) op(B_b) + beta D_b,   b = 0 .. batch-1
//
// column-major operands (element (i, j) of X at X[i + j ldx]), op = identity or
// transpose.  Used by the FP64 eigenpair refinement (A X, X^T [A X | X],
// X + X E), the blocked-reflector back-transformation and the warm-start
// products (Q0^T A Q0, Q0 V), and by the warm-start network's layers.
//
// Operand kinds (storage stays FP32 / FP64 in HBM; the matrix-core input type
// is chosen per call):
//   GK_F64   double in, v_mfma_f64_16x16x4f64, double accumulate
//   GK_F32   float in,  v_mfma_f32_16x16x4f32, float accumulate
//   GK_F16   float in rounded to half on the way into LDS, v_mfma_f32_16x16x16f16,
//            float accumulate
//   GK_BF16  float in rounded (nearest even) to bfloat16, v_mfma_f32_16x16x16bf16_1k,
//            float accumulate
//   GK_F16X3 float in split into two halves, x = hi + lo / 2^11 with hi = half(x) and
//            lo = half((x - hi) 2^11), and multiplied as hi*hi + (hi*lo + lo*hi) / 2^11 on
//            v_mfma_f32_16x16x16f16 (three products, two FP32 accumulators): FP16
//            matrix cores at close to FP32 accuracy (the error-compensated splitting
//            of Ootomo & Yokota, "Recovering single precision accuracy from Tensor
//            Cores while surpassing the FP32 theoretical peak performance", IJHPCA
//            2022); |x| must stay below 65504
//
// Shape: a 256-thread workgroup computes a 64 x 64 tile of one C_b; wave q owns
// the 32 x 32 quarter (q & 1, q >> 1), i.e. 2 x 2 MFMA tiles of 16 x 16.  The
// contraction runs in chunks of 16: a chunk's global loads are issued two
// chunks ahead, into one of two register stages, and stored to LDS (k-major, so
// that the lanes of an MFMA operand read consecutive words) one chunk ahead,
// into one of two LDS buffers while the other one is multiplied.  Out-of-range rows, columns
// and k are zero-filled, so any M, N, K >= 1 is valid.
//
// sym > 0: C_b is a row of `sym`-wide square blocks of tiles whose results are
// symmetric (X^T X, X^T A X); only tiles on or above each block's diagonal are
// computed and stored (the lower part of each block is left untouched).
//
// XCD locality: consecutive workgroup ids go to different XCDs (round robin over
// 8), so the ids are remapped to put consecutive tiles of one C_b (which share
// operand panels in L2) on the same XCD when the tile count allows it.
//
// MFMA operand layout (wave64, 16 x 16 output): lane l = m + 16 g holds
// A(m, k) and B(k, m) for k = g (x4 forms) or k = 4 g .. 4 g + 3 (x16 forms);
// accumulator slot r of lane l holds C(4 g + r, m), except FP64 (g + 4 r).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace eigmi {

enum GemmKind { GK_F64 = 0, GK_F32 = 1, GK_F16 = 2, GK_BF16 = 3, GK_F16X3 = 4 };
constexpr float GM_X3_SCALE = 2048.0f;   // 2^11: the low half's scale

struct GemmArgs {
  int M, N, K, batch;
  int transa, transb;
  const void *A; long long lda, sa;
  const void *B; long long ldb, sb;
  const void *D; long long ldd, sd;   // D may be null (beta ignored) and may equal C
  void *C; long long ldc, sc;
  double alpha, beta;
  int sym;                              // 0, or the width (in 64-wide tiles) of the symmetric blocks
};

constexpr int GM_TILE = 64, GM_KC = 16, GM_THREADS = 256, GM_PAD = 4;

template <int KIND> struct GemmT;
template <> struct GemmT<GK_F64> {
  typedef double S;       // storage type
  typedef double L;       // LDS / operand type
  typedef double Acc;
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ L cvt(S x) { return x; }
};
template <> struct GemmT<GK_F32> {
  typedef float S;
  typedef float L;
  typedef float Acc;
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ L cvt(S x) { return x; }
};
template <> struct GemmT<GK_F16> {
  typedef float S;
  typedef _Float16 L;
  typedef float Acc;
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ L cvt(S x) { return (_Float16)x; }
};
template <> struct GemmT<GK_BF16> {
  typedef float S;
  typedef unsigned short L;  // bfloat16 bits
  typedef float Acc;
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ L cvt(S x)
  {
    unsigned int u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest, ties to even
    return (unsigned short)(u >> 16);
  }
};

struct HalfPair {
  _Float16 hi, lo;
};
template <> struct GemmT<GK_F16X3> {
  typedef float S;
  typedef HalfPair L;
  typedef float Acc;
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ L cvt(S x)
  {
    L r;
    r.hi = (_Float16)x;
    r.lo = (_Float16)((x - (float)r.hi) * GM_X3_SCALE);
    return r;
  }
};

// one MFMA k-step group of a chunk for the 16 x 16 output tile
template <int KIND> struct GemmMma;
template <> struct GemmMma<GK_F64> {
  typedef GemmT<GK_F64>::v4 v4;
  // 4 k-steps of 4 per chunk of 16
  static __device__ __forceinline__ void chunk(const double (*As)[GM_TILE + GM_PAD],
                                               const double (*Bs)[GM_TILE + GM_PAD], int r0, int c0, int m,
                                               int g, v4 (&acc)[2][2])
  {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const double a0 = As[4 * ks + g][r0 + m], a1 = As[4 * ks + g][r0 + 16 + m];
      const double b0 = Bs[4 * ks + g][c0 + m], b1 = Bs[4 * ks + g][c0 + 16 + m];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
  static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};
template <> struct GemmMma<GK_F32> {
  typedef GemmT<GK_F32>::v4 v4;
  static __device__ __forceinline__ void chunk(const float (*As)[GM_TILE + GM_PAD],
                                               const float (*Bs)[GM_TILE + GM_PAD], int r0, int c0, int m, int g,
                                               v4 (&acc)[2][2])
  {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const float a0 = As[4 * ks + g][r0 + m], a1 = As[4 * ks + g][r0 + 16 + m];
      const float b0 = Bs[4 * ks + g][c0 + m], b1 = Bs[4 * ks + g][c0 + 16 + m];
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};
template <> struct GemmMma<GK_F16> {
  typedef GemmT<GK_F16>::v4 v4;
  typedef _Float16 h4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ void chunk(const _Float16 (*As)[GM_TILE + GM_PAD],
                                               const _Float16 (*Bs)[GM_TILE + GM_PAD], int r0, int c0, int m,
                                               int g, v4 (&acc)[2][2])
  {
    h4 a0, a1, b0, b1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a0[i] = As[4 * g + i][r0 + m];
      a1[i] = As[4 * g + i][r0 + 16 + m];
      b0[i] = Bs[4 * g + i][c0 + m];
      b1[i] = Bs[4 * g + i][c0 + 16 + m];
    }
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x16f16(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x16f16(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x16f16(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x16f16(a1, b1, acc[1][1], 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};
template <> struct GemmMma<GK_F16X3> {
  typedef GemmT<GK_F16X3>::v4 v4;
  typedef _Float16 h4 __attribute__((ext_vector_type(4)));
  // acc gets hi*hi, acc2 the cross terms (scaled by 2^11)
  static __device__ __forceinline__ void chunk(const HalfPair (*As)[GM_TILE + GM_PAD],
                                               const HalfPair (*Bs)[GM_TILE + GM_PAD], int r0, int c0, int m,
                                               int g, v4 (&acc)[2][2], v4 (&acc2)[2][2])
  {
    h4 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        const HalfPair a = As[4 * g + i][r0 + 16 * x + m], b = Bs[4 * g + i][c0 + 16 * x + m];
        ah[x][i] = a.hi;
        al[x][i] = a.lo;
        bh[x][i] = b.hi;
        bl[x][i] = b.lo;
      }
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah[x], bh[y], acc[x][y], 0, 0, 0);
        acc2[x][y] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah[x], bl[y], acc2[x][y], 0, 0, 0);
        acc2[x][y] = __builtin_amdgcn_mfma_f32_16x16x16f16(al[x], bh[y], acc2[x][y], 0, 0, 0);
      }
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};
template <> struct GemmMma<GK_BF16> {
  typedef GemmT<GK_BF16>::v4 v4;
  typedef short s4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ void chunk(const unsigned short (*As)[GM_TILE + GM_PAD],
                                               const unsigned short (*Bs)[GM_TILE + GM_PAD], int r0, int c0, int m,
                                               int g, v4 (&acc)[2][2])
  {
    s4 a0, a1, b0, b1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a0[i] = (short)As[4 * g + i][r0 + m];
      a1[i] = (short)As[4 * g + i][r0 + 16 + m];
      b0[i] = (short)Bs[4 * g + i][c0 + m];
      b1[i] = (short)Bs[4 * g + i][c0 + 16 + m];
    }
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, acc[1][1], 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

// Tile index of workgroup `bid` among `nt` tiles per product and `nb` products:
// consecutive tiles of one product on one XCD when the count divides evenly.
__device__ __forceinline__ long long gemm_xcd_remap(long long bid, long long total)
{
  if (total % 8 != 0) return bid;
  const long long per = total / 8;
  return (bid % 8) * per + bid / 8;
}

// Each thread loads 4 elements of the 64 x 16 A chunk and 4 of the 16 x 64 B
// chunk.  Non-transposed A (rows contiguous): thread t takes row t & 63 and k =
// (t >> 6) + 4 i; transposed A (k contiguous): k = t & 15, row = (t >> 4) + 16 i.
// B likewise with the roles of rows and k exchanged.
// held to 128 registers (four waves per SIMD): with the second chunk buffer
// the FP64 kind otherwise takes 136 and runs at three
template <int KIND>
__global__ __launch_bounds__(GM_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void eig_gemm_kernel(GemmArgs g)
{
  typedef GemmT<KIND> TT;
  typedef typename TT::S S;
  typedef typename TT::L Ls;
  typedef typename TT::v4 v4;
  // two chunk buffers: the next chunk is stored while the current one is
  // multiplied, with one barrier per chunk
  __shared__ Ls As[2][GM_KC][GM_TILE + GM_PAD];
  __shared__ Ls Bs[2][GM_KC][GM_TILE + GM_PAD];

  const int tm = (g.M + GM_TILE - 1) / GM_TILE, tn = (g.N + GM_TILE - 1) / GM_TILE;
  long long tiles;
  int symw = g.sym;
  if (symw > 0) {
    // tn = nblk * symw tiles in blocks of symw x symw (tm == symw)
    tiles = (long long)(tn / symw) * (symw * (symw + 1) / 2);
  } else {
    tiles = (long long)tm * tn;
  }
  const long long total = tiles * g.batch;
  const long long lid = gemm_xcd_remap(blockIdx.x, total);
  if (lid >= total) return;
  const long long b = lid / tiles;
  int t = (int)(lid - b * tiles), ti, tj;
  if (symw > 0) {
    const int per = symw * (symw + 1) / 2;
    const int blk = t / per;
    int r = t - blk * per, J = 0;
    while ((J + 1) * (J + 2) / 2 <= r) ++J;
    ti = r - J * (J + 1) / 2;
    tj = blk * symw + J;
  } else {
    ti = t % tm;
    tj = t / tm;
  }
  const int row0 = ti * GM_TILE, col0 = tj * GM_TILE;
  const S *__restrict__ A = (const S *)g.A + b * g.sa;
  const S *__restrict__ B = (const S *)g.B + b * g.sb;

  const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
  const int m = lane & 15, gg = lane >> 4;
  const int wr = (q & 1) * 32, wc = (q >> 1) * 32;

  // load geometry
  int ar[4], ak[4], bk[4], bc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (!g.transa) { ar[i] = tid & 63; ak[i] = (tid >> 6) + 4 * i; }
    else { ak[i] = tid & 15; ar[i] = (tid >> 4) + 16 * i; }
    if (!g.transb) { bk[i] = tid & 15; bc[i] = (tid >> 4) + 16 * i; }
    else { bc[i] = tid & 63; bk[i] = (tid >> 6) + 4 * i; }
  }
  // every load is issued, from indices clamped into the operand, and
  // out-of-range elements become zero when the chunk is stored to LDS: with no
  // branch or select after them the loads of the next chunks stay in flight
  // across this chunk's products
  const long long lsa_r = g.transa ? g.lda : 1, lsa_k = g.transa ? 1 : g.lda;
  const long long lsb_c = g.transb ? 1 : g.ldb, lsb_k = g.transb ? g.ldb : 1;
  auto loadA = [&](int k0, S (&ra)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = row0 + ar[i], k = k0 + ak[i];
      const int rc = r < g.M ? r : g.M - 1, kc = k < g.K ? k : g.K - 1;
      ra[i] = A[rc * lsa_r + kc * lsa_k];
    }
  };
  auto loadB = [&](int k0, S (&rb)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c = col0 + bc[i], k = k0 + bk[i];
      const int cc = c < g.N ? c : g.N - 1, kc = k < g.K ? k : g.K - 1;
      rb[i] = B[cc * lsb_c + kc * lsb_k];
    }
  };
  auto store = [&](int k0, const S (&ra)[4], const S (&rb)[4], Ls (*Ad)[GM_TILE + GM_PAD],
                   Ls (*Bd)[GM_TILE + GM_PAD]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool ina = row0 + ar[i] < g.M && k0 + ak[i] < g.K;
      const bool inb = col0 + bc[i] < g.N && k0 + bk[i] < g.K;
      Ad[ak[i]][ar[i]] = TT::cvt(ina ? ra[i] : S(0));
      Bd[bk[i]][bc[i]] = TT::cvt(inb ? rb[i] : S(0));
    }
  };

  v4 acc[2][2], acc2[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[x][y][r] = acc2[x][y][r] = 0;

  // two register stages: while chunk c is multiplied from LDS, chunk c + 1
  // (loaded during chunk c - 1) goes from registers into the other LDS buffer
  // and chunk c + 2 is loaded, so each global load has two chunks of products
  // to arrive in.  The loop body has no branch: loads past K read clamped
  // indices and stores past K write zeros (an odd chunk count multiplies one
  // zero chunk), which keeps the wait for a stage's loads from also waiting
  // for the other stage's
  S ra0[4], rb0[4], ra1[4], rb1[4];
  loadA(0, ra0);
  loadB(0, rb0);
  store(0, ra0, rb0, As[0], Bs[0]);
  loadA(GM_KC, ra1);
  loadB(GM_KC, rb1);
  __syncthreads();
  int cur = 0;
  // chunk k0 in LDS buffer cur, chunk k0 + GM_KC in (sa, sb), chunk k0 + 2 GM_KC
  // loaded into (la, lb), whose previous chunk is already in LDS
  auto step = [&](int k0, S (&la)[4], S (&lb)[4], const S (&sa)[4], const S (&sb)[4]) {
    loadA(k0 + 2 * GM_KC, la);
    loadB(k0 + 2 * GM_KC, lb);
    if constexpr (KIND == GK_F16X3) GemmMma<KIND>::chunk(As[cur], Bs[cur], wr, wc, m, gg, acc, acc2);
    else GemmMma<KIND>::chunk(As[cur], Bs[cur], wr, wc, m, gg, acc);
    // the buffer written here was last read before the previous barrier
    store(k0 + GM_KC, sa, sb, As[cur ^ 1], Bs[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  };
  for (int k0 = 0; k0 < g.K; k0 += 2 * GM_KC) {
    step(k0, ra0, rb0, ra1, rb1);
    step(k0 + GM_KC, ra1, rb1, ra0, rb0);
  }

  // epilogue: C = alpha acc + beta D
  typedef typename TT::Acc Acc;
  S *__restrict__ C = (S *)g.C + b * g.sc;
  const S *__restrict__ D = g.D ? (const S *)g.D + b * g.sd : nullptr;
  const Acc alpha = (Acc)g.alpha, beta = (Acc)g.beta;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = row0 + wr + 16 * x + GemmMma<KIND>::row(gg, r);
        const int j = col0 + wc + 16 * y + m;
        if (i < g.M && j < g.N) {
          Acc v = alpha * acc[x][y][r];
          if constexpr (KIND == GK_F16X3) v = alpha * (acc[x][y][r] + acc2[x][y][r] * (1.0f / GM_X3_SCALE));
          if (D) v += beta * (Acc)D[i + (size_t)j * g.ldd];
          C[i + (size_t)j * g.ldc] = (S)v;
        }
      }
}

} // namespace eigmi
