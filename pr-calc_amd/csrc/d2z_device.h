// Device helpers shared by the pass kernels: blocked layout, XCD tile
// mapping, the top-hat filter, the barrier and FP64 wave reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "d2z_kernels.h"

namespace prc {

// --------------------------------------------------------- blocked layout
// Complex half-spectrum arrays are stored as [kz/C][a][b][kz%C]: for every
// block of C consecutive kz the (a, b) plane is dense with C complex values per
// point.  A tile of C columns x all values of b is then one contiguous block
// (N*C*8 bytes), and a z row (a, b) is Nh/C segments of C*8 bytes.
//   spectrum S:        (a, b) = (ky, kx)
//   intermediates T:   (a, b) = (x, y)   after the first inverse pass
template <int N>
struct Blk {
  static constexpr int C = (N == 16 || N == 2048) ? 8 : 16;
};

// (kx, ky, kz) of blocked index id of the spectrum S = [kz/C][ky][kx][kz%C]
struct SIndex {
  int kx, ky, kz;
};
template <int C>
__device__ __forceinline__ SIndex s_index(long long id, int N) {
  SIndex s;
  const int c = (int)(id % C);
  const long long q = id / C;
  s.kx = (int)(q % N);
  const long long q2 = q / N;
  s.ky = (int)(q2 % N);
  s.kz = (int)(q2 / N) * C + c;
  return s;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Remap so each XCD works on a contiguous range of tiles: neighbouring tiles
// share pages and 128 B lines in the transposed accesses.
__device__ __forceinline__ long long xcd_tile(long long b, long long nblk) {
  if (nblk % 8 != 0) return b;
  return (b % 8) * (nblk / 8) + b / 8;
}

// ---------------------------------------------------- top-hat filter W(kR)
// sin and cos for 0 <= x < 2^12: Cody-Waite reduction by pi/2 in three parts
// with FMA, then the single-precision minimax polynomials of the Cephes
// library (S. L. Moshier) on [-pi/4, pi/4].
__device__ __forceinline__ void sincos_reduced(float x, float& s, float& c) {
  const float n = rintf(x * 0.636619772367581343f);
  float r = fmaf(-n, 1.5703125f, x);
  r = fmaf(-n, 4.837512969970703125e-4f, r);
  r = fmaf(-n, 7.54978995489188216e-8f, r);
  const float r2 = r * r;
  float ps = fmaf(r2, -1.9515295891e-4f, 8.3321608736e-3f);
  ps = fmaf(r2, ps, -1.6666654611e-1f);
  ps = fmaf(r2 * r, ps, r);
  float pc = fmaf(r2, 2.443315711809948e-5f, -1.388731625493765e-3f);
  pc = fmaf(r2, pc, 4.166664568298827e-2f);
  pc = fmaf(r2 * r2, pc, fmaf(-0.5f, r2, 1.0f));
  const int q = (int)n & 3;
  const float sv = (q & 1) ? pc : ps;
  const float cv = (q & 1) ? ps : pc;
  s = (q & 2) ? -sv : sv;
  c = ((q + 1) & 2) ? -cv : cv;
}

// W(x) = 3 (sin x - x cos x) / x^3; Taylor series below x = 1 where the closed
// form cancels (terms to x^10, truncation < 4e-11 on [0,1)).
__device__ __forceinline__ float tophat(float x) {
  if (x < 1.0f) {
    const float x2 = x * x;
    return 1.0f - x2 * (1.0f / 10.0f -
                        x2 * (1.0f / 280.0f -
                              x2 * (1.0f / 15120.0f - x2 * (1.0f / 1330560.0f - x2 * (1.0f / 172972800.0f)))));
  }
  float s, c;
  sincos_reduced(x, s, c);
  return 3.0f * fmaf(-x, c, s) * __builtin_amdgcn_rcpf(x * x * x);
}

// signed frequency index of grid index i (i >= N/2 maps to i - N), squared
__device__ __forceinline__ int fsq(int i, int N) {
  const int f = (i >= (N >> 1)) ? i - N : i;
  return f * f;
}

// W(|k| R) / N^3 for integer |k|^2 = m dk^2; k = 0 gives 0 (the reference
// zeroes the mean mode, delta2zreion_d2z.C:79-85).
__device__ __forceinline__ float filter_coef(int m, float dkR, float invN3) {
  if (m == 0) return 0.0f;
  return tophat(__builtin_amdgcn_sqrtf((float)m) * dkR) * invN3;
}

// W(|k| R) / N^3 from the interpolation table (entries c0..c3 per interval of
// length 1/invh in x = |k| R) for |k| = km dk; dkRh = dk R invh.  M = 0 or
// km = 0 fall back to the exact value / the zeroed mean mode.
__device__ __forceinline__ float filter_from_table(float km, float dkR, float dkRh, float invN3,
                                                   const float4* tab, int M) {
  if (km == 0.0f) return 0.0f;
  if (M == 0) return tophat(km * dkR) * invN3;
  const float u = km * dkRh;
  int i = (int)u;
  i = i < M - 1 ? i : M - 1;
  const float t = u - (float)i;
  const float4 c = tab[i];
  return fmaf(t, fmaf(t, fmaf(t, c.w, c.z), c.y), c.x) * invN3;
}

// The same in two steps for tables that reuse |k| over radii:
// kmag_grid(m) = |k| / dk, filter_of_kmag(kmag_grid(m), dkR, invN3) ==
// filter_coef(m, dkR, invN3) bit for bit.
__device__ __forceinline__ float kmag_grid(int m) { return __builtin_amdgcn_sqrtf((float)m); }
__device__ __forceinline__ float filter_of_kmag(float km, float dkR, float invN3) {
  if (km == 0.0f) return 0.0f;
  return tophat(km * dkR) * invN3;
}

// -------------------------------------- barrier: d2z() in double, stored float
// z = A*(delta_R + B) - 1 with A float and B = sqrt(2 (smin^2 - s^2)) * erfizeta
// in double (delta2zreion_d2z.C:206-216); separate multiply and subtract (no
// FMA), as the reference evaluates it.
__device__ __forceinline__ float barrier_z(float d, double A, double B) {
#pragma clang fp contract(off)
  const double s = (double)d + B;
  const double p = A * s;
  return (float)(p - 1.0);
}

// barrier_z of d * sc for a power-of-two sc: (double)d * sc is exact, so the
// FMA rounds once, as (double)(d * sc) + B does; sc = 1 gives barrier_z(d).
__device__ __forceinline__ float barrier_z_scaled(float d, double sc, double A, double B) {
#pragma clang fp contract(off)
  const double s = __builtin_fma((double)d, sc, B);
  const double p = A * s;
  return (float)(p - 1.0);
}

// FP32 form for the mixed-precision path, one FMA per cell: z = A' d + B' with
// A' = A sc (exact: A is a float and sc a power of two) and B' = A B - 1
// rounded once from double; |error| <= (|z| + |B'|) 2^-24 (~2e-6 for these
// barriers), far below that path's own FP16 error
struct BarrierF32 {
  float a, b;
};
__device__ __forceinline__ BarrierF32 barrier_f32(double A, double B, float sc) {
  return BarrierF32{(float)(A * (double)sc), (float)(A * B - 1.0)};
}
__device__ __forceinline__ float barrier_z_f32(float d, BarrierF32 k) { return fmaf(k.a, d, k.b); }

// Running maximum of barrier values for zreion, which starts at 0 and is only
// raised: values <= 0 never change it, so the maximum may start at 0, and for
// non-negative floats the signed-integer order of the bit patterns is the float
// order (a negative z has the sign bit set and compares below all of them).  One
// integer max, where fmaxf of a loop-carried value costs a canonicalizing max too.
__device__ __forceinline__ float zmax_nonneg(float best, float z) {
  return __int_as_float(max(__float_as_int(best), __float_as_int(z)));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace prc
