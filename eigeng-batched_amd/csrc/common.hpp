// common.hpp - device helpers for the gfx950 batched eigensolver kernels.
//
// Group collectives: a "group" is NB consecutive lanes of one wave64 (NB = 8,
// 16, 32 or 64) that own one matrix.  Reductions use DPP row operations for the
// lanes of a row of 16, ds_swizzle for 16 <-> 16 inside a half wave and
// ds_bpermute for the two halves; every step is a symmetric butterfly so all
// lanes of the group end with bitwise identical results (group-uniform values
// must be exactly uniform because they steer branches).
//
// Reference counterparts: the cooperative-groups butterflies sum_on_cg /
// min_on_cg / bcast_on_cg (misc_gpu.hpp:614-786) and pythag/pythag1
// (misc_gpu.hpp:514-548).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace eigmi {

template <typename T> struct vec2;
template <> struct vec2<float> { typedef float2 type; };
template <> struct vec2<double> { typedef double2 type; };

template <typename T> __device__ __forceinline__ T eps_of();
template <> __device__ __forceinline__ float eps_of<float>() { return 1.1920928955078125e-07f; }
template <> __device__ __forceinline__ double eps_of<double>() { return 2.220446049250313e-16; }

// Ordering point for LDS traffic between lanes of one wave: LDS operations of a
// wave execute in order, so only compiler code motion has to be stopped.
__device__ __forceinline__ void wsync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- raw 32-bit lane movement -------------------------------------------
template <int CTRL>
__device__ __forceinline__ int dpp32(int x)
{
  return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ int swz_xor16(int x) { return __builtin_amdgcn_ds_swizzle(x, 0x401F); }
__device__ __forceinline__ int bperm_xor32(int x)
{
  const int src = (int)((threadIdx.x ^ 32u) & 63u) << 2;
  return __builtin_amdgcn_ds_bpermute(src, x);
}

// Apply a 32-bit lane movement to float / double / int.
template <typename F>
__device__ __forceinline__ float lane_move(float x, F f) { return __int_as_float(f(__float_as_int(x))); }
template <typename F>
__device__ __forceinline__ int lane_move(int x, F f) { return f(x); }
template <typename F>
__device__ __forceinline__ double lane_move(double x, F f)
{
  const long long b = __double_as_longlong(x);
  const int lo = f((int)(b & 0xffffffffll));
  const int hi = f((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// quad_perm [1,0,3,2] = xor 1; quad_perm [2,3,0,1] = xor 2;
// row_half_mirror pairs the two quads of 8 lanes; row_mirror pairs the halves of 16.
struct MvX1 { __device__ int operator()(int v) const { return dpp32<0xB1>(v); } };
struct MvX2 { __device__ int operator()(int v) const { return dpp32<0x4E>(v); } };
struct MvHM { __device__ int operator()(int v) const { return dpp32<0x141>(v); } };
struct MvRM { __device__ int operator()(int v) const { return dpp32<0x140>(v); } };
struct MvX16 { __device__ int operator()(int v) const { return swz_xor16(v); } };
struct MvX32 { __device__ int operator()(int v) const { return bperm_xor32(v); } };

// Butterfly all-reduce over the NB lanes of a group with a commutative op.
template <int NB, typename T, typename Op>
__device__ __forceinline__ T group_reduce(T x, Op op)
{
  static_assert(NB == 8 || NB == 16 || NB == 32 || NB == 64, "group size");
  x = op(x, lane_move(x, MvX1()));
  x = op(x, lane_move(x, MvX2()));
  x = op(x, lane_move(x, MvHM()));
  if (NB >= 16) x = op(x, lane_move(x, MvRM()));
  if (NB >= 32) x = op(x, lane_move(x, MvX16()));
  if (NB >= 64) x = op(x, lane_move(x, MvX32()));
  return x;
}
struct OpAdd { template <typename T> __device__ T operator()(T a, T b) const { return a + b; } };

template <int NB, typename T>
__device__ __forceinline__ T group_sum(T x) { return group_reduce<NB>(x, OpAdd()); }

// Broadcast from lane `src` (0..63, absolute) to the calling lanes.
template <typename T>
__device__ __forceinline__ T lane_bcast(T x, int src)
{
  const int addr = (src & 63) << 2;
  return lane_move(x, [addr](int v) { return __builtin_amdgcn_ds_bpermute(addr, v); });
}

// Ballot restricted to this lane's group, shifted so bit 0 is the group's lane 0.
template <int NB>
__device__ __forceinline__ uint64_t group_ballot(bool p)
{
  const uint64_t all = __ballot(p);
  if (NB == 64) return all;
  const int base = (int)(threadIdx.x & 63u) & ~(NB - 1);
  return (all >> base) & ((1ull << NB) - 1ull);
}

// ---- scalar arithmetic ---------------------------------------------------
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float sqrt_(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float abs_(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double abs_(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float sign_(float x, float y) { return __builtin_copysignf(x, y); }
__device__ __forceinline__ double sign_(double x, double y) { return __builtin_copysign(x, y); }
template <typename T> __device__ __forceinline__ T nz(T x) { return x == T(0) ? T(1) : x; }

// sqrt(x^2 + y^2) with one compensated correction (Borges, ACM TOMS 2020):
// the rounding errors of the two squares and of their sum are folded back
// into the square root, giving a nearly correctly rounded result without
// scaling (same method as the reference's pythag, misc_gpu.hpp:533-548).
template <typename T>
__device__ __forceinline__ T hypot_c(T x, T y)
{
  const T xx = x * x, yy = y * y;
  const T sig = xx + yy;
  const T h = sqrt_(sig);
  const T sig_err = (sig - xx) - yy;
  const T tau = fma_(y, y, -yy) + fma_(x, x, -xx) - sig_err + fma_(-h, h, sig);
  return fma_(tau / nz(h), T(0.5), h);
}
template <typename T>
__device__ __forceinline__ T hypot1_c(T x) { return hypot_c(x, T(1)); }

// 1/sqrt(x) from the hardware estimate plus Newton steps y <- y*(3 - x*y*y)/2:
// one step for float (estimate ~1 ulp -> ~0.5 ulp + rounding), two for double
// (estimate ~2^-22 -> full precision).  x must be positive and normal.
__device__ __forceinline__ float rsqrt_nr(float x)
{
  float y = __builtin_amdgcn_rsqf(x);
  const float hx = 0.5f * x;
  y = y * __builtin_fmaf(-hx * y, y, 1.5f);
  return y;
}
__device__ __forceinline__ double rsqrt_nr(double x)
{
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = y * __builtin_fma(-hx * y, y, 1.5);
  y = y * __builtin_fma(-hx * y, y, 1.5);
  return y;
}

// Givens parameters for (f, g): r = sqrt(f^2 + g^2), s = f/r, c = g/r.
// Returns false (and r = 0) when f = g = 0.  The float version renormalises
// (c, s) once with the fma-exact residual e = c^2 + s^2 - 1, so c^2 + s^2 - 1
// is O(eps^2) plus final rounding: no bias accumulates in the eigenvectors.
__device__ __forceinline__ bool givens(float f, float g, float &r, float &c, float &s)
{
  const float ss = __builtin_fmaf(f, f, g * g);
  if (ss == 0.0f) { r = 0.0f; return false; }
  const float ri = rsqrt_nr(ss);
  r = ss * ri;
  c = g * ri;
  s = f * ri;
  const float h = -0.5f * __builtin_fmaf(c, c, __builtin_fmaf(s, s, -1.0f));
  c = __builtin_fmaf(c, h, c);
  s = __builtin_fmaf(s, h, s);
  return true;
}
__device__ __forceinline__ bool givens(double f, double g, double &r, double &c, double &s)
{
  const double ss = __builtin_fma(f, f, g * g);
  if (ss == 0.0) { r = 0.0; return false; }
  const double ri = rsqrt_nr(ss);
  r = ss * ri;
  c = g * ri;
  s = f * ri;
  return true;
}

// givens() for a known ss = f^2 + g^2 > 0 (the QL steps' form): the same
// values, computed without the zero test.
__device__ __forceinline__ void givens_nz(float f, float g, float ss, float &r, float &c, float &s)
{
  const float ri = rsqrt_nr(ss);
  const float cn = g * ri, sn = f * ri;
  const float h = -0.5f * __builtin_fmaf(cn, cn, __builtin_fmaf(sn, sn, -1.0f));
  r = ss * ri;
  c = __builtin_fmaf(cn, h, cn);
  s = __builtin_fmaf(sn, h, sn);
}
__device__ __forceinline__ void givens_nz(double f, double g, double ss, double &r, double &c, double &s)
{
  const double ri = rsqrt_nr(ss);
  r = ss * ri;
  c = g * ri;
  s = f * ri;
}

// ---- one value per lane of a group: element i lives in lane i ------------
// Group-uniform index i.  For a full wave the element is read with
// v_readlane (the index is wave-uniform, so it sits in an SGPR); for smaller
// groups with ds_bpermute.
template <int NB>
__device__ __forceinline__ float gread(float v, int i)
{
  if (NB == 64) return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), i));
  const int base = (int)(threadIdx.x & 63u) & ~(NB - 1);
  return __int_as_float(__builtin_amdgcn_ds_bpermute((base + i) << 2, __float_as_int(v)));
}
template <int NB>
__device__ __forceinline__ int gread(int v, int i)
{
  if (NB == 64) return __builtin_amdgcn_readlane(v, i);
  const int base = (int)(threadIdx.x & 63u) & ~(NB - 1);
  return __builtin_amdgcn_ds_bpermute((base + i) << 2, v);
}
template <int NB>
__device__ __forceinline__ double gread(double v, int i)
{
  const long long b = __double_as_longlong(v);
  const int lo = gread<NB>((int)(b & 0xffffffffll), i);
  const int hi = gread<NB>((int)(b >> 32), i);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 8 consecutive T from LDS with 16-byte loads (p must be 16-byte aligned).
__device__ __forceinline__ void ld8(const float *p, float o[8])
{
  const float4 a = *(const float4 *)p, b = *(const float4 *)(p + 4);
  o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w; o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
}
__device__ __forceinline__ void ld8(const double *p, double o[8])
{
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double2 v = *(const double2 *)(p + 2 * q);
    o[2 * q] = v.x;
    o[2 * q + 1] = v.y;
  }
}
__device__ __forceinline__ void st8(float *p, const float o[8])
{
  *(float4 *)p = make_float4(o[0], o[1], o[2], o[3]);
  *(float4 *)(p + 4) = make_float4(o[4], o[5], o[6], o[7]);
}
__device__ __forceinline__ void st8(double *p, const double o[8])
{
#pragma unroll
  for (int q = 0; q < 4; ++q) *(double2 *)(p + 2 * q) = make_double2(o[2 * q], o[2 * q + 1]);
}

// One step of a butterfly reduce-scatter: a lane and its partner under Mv
// hold 2K partial values for the same 2K indices and opposite `side`s; the lane
// keeps half `side` (indices [K side, K side + K)) with the partner's partials
// added, in v[0..K).
template <int K, typename T, typename Mv>
__device__ __forceinline__ void reduce_scatter_step(T *v, bool side, Mv mv)
{
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const T keep = side ? v[i + K] : v[i];
    const T send = side ? v[i] : v[i + K];
    v[i] = keep + lane_move(send, mv);
  }
}

// Pair sums over the gfx950 lane swaps (VALU, no LDS): v_permlane16_swap
// exchanges the odd 16-lane row of each pair (rows 2k, 2k+1) of its first
// operand with the even row of its second; v_permlane32_swap the upper 32
// lanes of the first with the lower 32 of the second.  With (a, b) in every
// lane, swap_sum<W>(a, b) returns on the low side (even row / lanes 0-31)
// a + the partner lane's a, and on the high side b + the partner's b, where
// the partner is lane ^ W: one butterfly reduce-scatter step of one pair.
template <int W>
__device__ __forceinline__ unsigned swap_u32(unsigned a, unsigned b, unsigned &bo)
{
  static_assert(W == 16 || W == 32, "row pairs or wave halves");
  if constexpr (W == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
    bo = r[1];
    return r[0];
  } else {
    const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    bo = r[1];
    return r[0];
  }
}
template <int W>
__device__ __forceinline__ float swap_sum(float a, float b)
{
  unsigned bo;
  const unsigned ao = swap_u32<W>(__float_as_uint(a), __float_as_uint(b), bo);
  return __uint_as_float(ao) + __uint_as_float(bo);
}
template <int W>
__device__ __forceinline__ double swap_sum(double a, double b)
{
  const unsigned long long ab = (unsigned long long)__double_as_longlong(a);
  const unsigned long long bb = (unsigned long long)__double_as_longlong(b);
  unsigned blo, bhi;
  const unsigned alo = swap_u32<W>((unsigned)ab, (unsigned)bb, blo);
  const unsigned ahi = swap_u32<W>((unsigned)(ab >> 32), (unsigned)(bb >> 32), bhi);
  return __longlong_as_double((long long)(((unsigned long long)ahi << 32) | alo)) +
         __longlong_as_double((long long)(((unsigned long long)bhi << 32) | blo));
}

} // namespace eigmi
