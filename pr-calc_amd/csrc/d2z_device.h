// Device helpers shared by the pass kernels: blocked layout, XCD tile
// mapping, the top-hat filter, the barrier and FP64 wave reductions.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
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

// Tile of block b for the x passes, whose tiles are numbered so that each
// run of `group` consecutive tiles writes whole 128 B lines of the output
// together (the column parts of a row, and the adjacent rows sharing a line
// when rows are shorter than 128 B): a group runs on one XCD (its L2 merges
// the partial lines) and consecutive groups on consecutive XCDs (blocks go to
// the XCDs round-robin), so the resident blocks of all XCDs write one run of
// consecutive lines of each output row.  For FP32 outputs this measured
// faster than xcd_tile (each XCD on its own range of tiles), same bits: at
// 1024^3 x 96 the split x pass 122-136 against 137-145 ms per step, the FP32
// vector x pass 146-162 against 167-169, at 2048^3 x 128 2074 against 2136.
// The FP16-output x passes measured slower with it (at 1024^3 98-106 against
// 97-98 ms, at 2048^3 1871 against 1769) and keep xcd_tile.
__device__ __forceinline__ long long line_tile(long long b, long long nblk, int group) {
  if (nblk % (8LL * group) != 0) return b;
  const long long slot = b >> 3;
  return ((slot / group) * 8 + (b & 7)) * group + slot % group;
}

// Tiles per 128 B output line: nsub column parts of rows of row_bytes.
constexpr int line_group(int nsub, int row_bytes) { return row_bytes >= 128 ? nsub : nsub * (128 / row_bytes); }

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

// One radius's W table from a |k| table: wt[q] = filter_from_table(kt[q], ...)
// for q < n, ITERS strided entries per thread of NT (ITERS NT >= n).  With an
// interpolation table (M > 0) a thread's entries are independent and written
// without branches, so their LDS reads overlap instead of queueing one after
// the other; the values are filter_from_table's bit for bit.
template <int NT, int ITERS>
__device__ __forceinline__ void fill_wtable(float* wt, const float* kt, int n, float dkR, float dkRh, float scale,
                                            const float4* tab, int M, int tid) {
  if (ITERS <= 8 && M > 0) {
#pragma unroll
    for (int k = 0; k < ITERS; ++k) {
      const int q = tid + k * NT;
      if (q < n) {
        const float km = kt[q];
        const float u = km * dkRh;
        int i = (int)u;
        i = i < M - 1 ? i : M - 1;
        const float t = u - (float)i;
        const float4 c = tab[i];
        const float w = fmaf(t, fmaf(t, fmaf(t, c.w, c.z), c.y), c.x) * scale;
        wt[q] = km == 0.0f ? 0.0f : w;
      }
    }
  } else {
    for (int q = tid; q < n; q += NT) wt[q] = filter_from_table(kt[q], dkR, dkRh, scale, tab, M);
  }
}

// The same in two steps for tables that reuse |k| over radii:
// kmag_grid(m) = |k| / dk, filter_of_kmag(kmag_grid(m), dkR, invN3) ==
// filter_coef(m, dkR, invN3) bit for bit.
__device__ __forceinline__ float kmag_grid(int m) { return __builtin_amdgcn_sqrtf((float)m); }
__device__ __forceinline__ float filter_of_kmag(float km, float dkR, float invN3) {
  if (km == 0.0f) return 0.0f;
  return tophat(km * dkR) * invN3;
}

// ------------------------------------------- per-radius barrier offset B_R
// B = sqrt(2 (smin^2 - sigma^2)) * erfizeta with the float difference widened
// to double (delta2zreion_d2z.C:211).  Where sigma(R) > sigma_min the
// reference's square root is NaN, its z is NaN and its strict `zcur >
// zreion` never updates (delta2zreion_d2z.C:117,138): such a radius
// contributes nothing.  Here B is -inf instead, so z = -inf, which no form of
// the running maximum (strict compare, fmaxf, or the integer max of
// zmax_nonneg, where a positive NaN would win) lets through.
__host__ __device__ inline double barrier_offset(float dd, float erfizeta) {
  if (!(dd >= 0.0f)) return -INFINITY;
  return sqrt(2. * (double)dd) * (double)erfizeta;
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
// order (a negative z, -inf included, has the sign bit set and compares below
// all of them).  z must not be NaN (a positive NaN's bits exceed every finite
// float's): barrier_offset makes the radii whose reference z is NaN give -inf.
// One integer max, where fmaxf of a loop-carried value costs a canonicalizing
// max too.
__device__ __forceinline__ float zmax_nonneg(float best, float z) {
  return __int_as_float(max(__float_as_int(best), __float_as_int(z)));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// The same in float: the FP16-storage paths sum their per-radius squares over
// a wave in float (64 lanes; relative error far below their own) and widen to
// double for the sums over waves and blocks
__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ------------------------------------------------------ FP16 complex pairs
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pack_h2(float re, float im) {
  const __half2 h = __floats2half2_rn(re, im);
  return *reinterpret_cast<const uint32_t*>(&h);
}
__device__ __forceinline__ float2 unpack_h2(uint32_t u) {
  __half2 h;
  *reinterpret_cast<uint32_t*>(&h) = u;
  return __half22float2(h);
}

// Real-output glue of the half-length c2r in packed FP16: Z[k] = P a + Q conj(m)
// with a = X[k], m = X[(L - k) mod L], P = 1 + i w^k, Q = 1 - i w^k (w = exp(2 pi
// i / N)), the same as (a + conj m) + i w^k (a - conj m); P = 0, Q = 1 + i at
// k = 0 give the packed slot's (a.x + a.y, a.x - a.y).  Four packed FMAs, the
// swaps and signs in their operand modifiers.
__device__ __forceinline__ uint32_t c2r_glue_h2(uint32_t av, uint32_t mv, uint32_t pv, uint32_t qv) {
  const half2_t a = __builtin_bit_cast(half2_t, av), m = __builtin_bit_cast(half2_t, mv);
  const half2_t P = __builtin_bit_cast(half2_t, pv), Q = __builtin_bit_cast(half2_t, qv);
  half2_t z = a * P.xx;                                 // (P.x a.x, P.x a.y)
  z = __builtin_elementwise_fma(a.yx, (half2_t){-P.y, P.y}, z);  // (-P.y a.y, P.y a.x)
  z = __builtin_elementwise_fma(m, (half2_t){Q.x, -Q.x}, z);     // (Q.x m.x, -Q.x m.y)
  z = __builtin_elementwise_fma(m.yx, Q.yy, z);                  // (Q.y m.y, Q.y m.x)
  return __builtin_bit_cast(uint32_t, z);
}
// The glue constants (P, Q) of c2r_glue_h2 for k, from w^k = exp(2 pi i k/N)
__device__ __forceinline__ uint2 c2r_glue_pq(int k, float2 wk) {
  return k == 0 ? make_uint2(pack_h2(0.f, 0.f), pack_h2(1.f, 1.f))
                : make_uint2(pack_h2(1.f - wk.y, wk.x), pack_h2(1.f + wk.y, -wk.x));
}

// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): counter c, key k, ten rounds
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k.x += W0;
      k.y += W1;
    }
    const uint64_t p0 = (uint64_t)M0 * c.x;
    const uint64_t p1 = (uint64_t)M1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
  }
  return c;
}

}  // namespace prc
