// Matrix-core (MFMA) variants of the fused x and z passes for the mixed path
// at N = 1024 (FP16 intermediates).  x pass: the two radix-16 stages of the
// 1024-point inverse transform (1024 = 16 * 16 * 4) run as 32 x 32 real matrix
// products on v_mfma_f32_16x16x32_f16 (FP16 operands, FP32 accumulation), the
// radix-4 stage on the vector ALUs.  Reference: delta2zreion_d2z.C:58-92
// (filter and c2r), with the same W(kR) table and the same output layout as
// k_fx.  The z pass (below) does the same for its 512-point half-length c2r.
//
// Stockham formulation (as fft_core.h): stage (p, R) takes butterfly b's
// inputs x[b + r L/R], multiplies input r by w^(r (b mod p)) with w =
// exp(+2 pi i / (p R)), applies the R-point DFT and writes output q to
// y[(b/p) p R + (b mod p) + q p].
//
// One MFMA computes a radix-16 DFT of 16 butterflies with the same k = b mod p
// (the twiddles are folded into that k's matrix G_k): B (32 x 16) holds the
// inputs, K = 8 g + 2 j + c <-> input r = 4 g + j (c = re/im), column n =
// butterfly; A (32 x 32 as two 16-row blocks h) holds G_k with row
// M = 16 h + 4 g + 2 s + c <-> output q = g + 4 s + 8 h.  Lane l = 16 g + n
// supplies B's rows 8g..8g+7 of column n (4 complex inputs) and receives D's
// rows 4g..4g+3 of column n for each h: outputs q = g, g+4 (h = 0) and g+8,
// g+12 (h = 1).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d2z_device.h"
#include "d2z_kernels.h"
#include "fft_core.h"

namespace prc {
namespace {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kN = 1024;         // transform length and grid size
constexpr int kC = 16;           // blocked-layout width (Blk<1024>::C)
constexpr int kCF = 8;           // columns (kz) per block
constexpr int kNSUB = kC / kCF;  // column sub-tiles per kz block
constexpr int kNF = kN / 2 + 1;  // distinct |fx|
constexpr int kThreads = 1024;   // 16 waves

// One 16-row block (h) of the matrix G_k of a radix-16 stage with p, in the
// A-operand layout of lane l: A[m = l % 16][K = 8 (l / 16) + j], j < 8.
__device__ half8_t stage_matrix(int k, int p, int h, int l) {
  const int m = l & 15, gp = l >> 4;
  const int g = m >> 2, s = (m >> 1) & 1, c = m & 1;
  const int q = g + 4 * s + 8 * h;
  half8_t a;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int r = 4 * gp + (j >> 1), cp = j & 1;
    // angle 2 pi (q r / 16 + r k / (16 p)) = pi * theta / (8 p)
    const int theta = (p * q * r + r * k) % (16 * p);
    float sn, cs;
    sincospif((float)theta / (float)(8 * p), &sn, &cs);
    float v;
    if (c == 0) v = cp == 0 ? cs : -sn;
    else v = cp == 0 ? sn : cs;
    a[j] = (_Float16)v;
  }
  return a;
}

// LDS exchange buffers: 1024 rows of kCF complex FP16 (one dword each); row
// pos is stored at pos ^ (bit 8 of pos | bit 4 of pos << 1), which spreads
// every access pattern of the three stages over all 32 banks.
__device__ __forceinline__ int xrow(int pos) { return pos ^ (((pos >> 8) & 1) | (((pos >> 4) & 1) << 1)); }

// B operand from 4 complex values
__device__ __forceinline__ half8_t pack_b(const float2 (&x)[4]) {
  half8_t b;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    b[2 * j] = (_Float16)x[j].x;
    b[2 * j + 1] = (_Float16)x[j].y;
  }
  return b;
}
// B operand from 4 stored complex FP16 values (low half = re, high = im)
__device__ __forceinline__ half8_t load_b(const uint32_t* buf, int row0, int rstride, int c) {
  union {
    uint32_t u[4];
    half8_t h;
  } cv;
#pragma unroll
  for (int j = 0; j < 4; ++j) cv.u[j] = buf[xrow(row0 + j * rstride) * kCF + c];
  return cv.h;
}

// D outputs of the two blocks -> 4 complex values q = g, g+4, g+8, g+12
__device__ __forceinline__ void store_d(uint32_t* buf, const float4_t& d0, const float4_t& d1, int row_q0,
                                        int qstride, int c) {
  buf[xrow(row_q0) * kCF + c] = pack_h2(d0[0], d0[1]);
  buf[xrow(row_q0 + 4 * qstride) * kCF + c] = pack_h2(d0[2], d0[3]);
  buf[xrow(row_q0 + 8 * qstride) * kCF + c] = pack_h2(d1[0], d1[1]);
  buf[xrow(row_q0 + 12 * qstride) * kCF + c] = pack_h2(d1[2], d1[3]);
}

// Operand table: set 0, 1 = stage (k = 0, p = 1) blocks h = 0, 1; set
// 2 + 2 k + h = stage (k, p = 16) block h; [set][lane], built once at set-up
constexpr int kMatSets = 2 + 2 * 16;
__device__ __forceinline__ int mat_set_a(int h) { return h; }
__device__ __forceinline__ int mat_set_b(int k, int h) { return 2 + 2 * k + h; }

__global__ void __launch_bounds__(64) k_mfma_mats(half8_t* __restrict__ out) {
  const int set = blockIdx.x, l = threadIdx.x;
  out[set * 64 + l] = set < 2 ? stage_matrix(0, 1, set, l) : stage_matrix((set - 2) >> 1, 16, (set - 2) & 1, l);
}

__global__ void __launch_bounds__(kThreads)
    k_fx_mfma(const float2* __restrict__ S, const float2* __restrict__ D0, const float2* __restrict__ DN,
              __half2* __restrict__ Tbuf, float dk, float invN3, const RadiusConst* __restrict__ rc, int r0, int nb,
              const float2* __restrict__ tw_full, const float4* __restrict__ wcoef, int wM, float winvh,
              const half8_t* __restrict__ mats) {
  constexpr long long Nh = kN / 2;
  constexpr long long plane = (long long)kN * kN * Nh;
  __shared__ uint32_t bufX[kN * kCF];
  __shared__ uint32_t bufY[kN * kCF];
  __shared__ float wtab[(kCF + 1) * kNF];  // [|fx|][column], then the kz = N/2 column
  __shared__ float ktab[(kCF + 1) * kNF];  // |k|/dk of the same entries
  extern __shared__ float4 wsh[];           // W(x) interpolation table (wM intervals)
  float* wpk = wtab + kCF * kNF;

  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int g = l >> 4, n = l & 15;
  const int c = n & 7, sel = n >> 3;  // MFMA column n = (column c, butterfly sel)

  const long long tile = xcd_tile(blockIdx.x, gridDim.x);
  const int sub = (int)(tile % kNSUB);
  const int ky = (int)((tile / kNSUB) % kN);
  const int kzb = (int)(tile / (kNSUB * kN));
  const int kz0 = kzb * kC + sub * kCF;
  const bool has_packed = (kzb == 0 && sub == 0);  // block-uniform
  const bool packed = has_packed && c == 0;        // this lane's column is kz = 0

  // matrices: stage A (p = 1, k = 0) and stage B (p = 16, k = this wave)
  const half8_t gA0 = mats[mat_set_a(0) * 64 + l], gA1 = mats[mat_set_a(1) * 64 + l];
  const half8_t gB0 = mats[mat_set_b(w, 0) * 64 + l], gB1 = mats[mat_set_b(w, 1) * 64 + l];

  // the lane's 8 spectrum values: stage-A butterflies b = 4w + 2t + sel,
  // inputs r = 4g + j at kx = b + 64 r
  const float2* sp = S + ((long long)kzb * kN + ky) * kN * kC + sub * kCF + c;
  float2 m[8];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kx = 4 * w + 2 * t + sel + 64 * (4 * g + j);
      m[4 * t + j] = packed ? D0[(long long)kx * kN + ky] : sp[(long long)kx * kC];
    }
  }
  const int fy2 = fsq(ky, kN);
  const int ntab = (has_packed ? kCF + 1 : kCF) * kNF;
  for (int q = tid; q < ntab; q += kThreads) {
    int f, kzq;
    if (q < kCF * kNF) {
      f = q / kCF;
      kzq = kz0 + q % kCF;
    } else {
      f = q - kCF * kNF;
      kzq = kN / 2;
    }
    ktab[q] = kmag_grid(f * f + fy2 + kzq * kzq);
  }
  for (int q = tid; q < wM; q += kThreads) wsh[q] = wcoef[q];
  // stage-C thread layout: column cl, butterflies i and i + 128
  const int cl = tid & 7, ic = tid >> 3;
  __half2* tp = Tbuf + (long long)kzb * kN * kN * kC + (long long)ky * kC + sub * kCF + cl;
  __syncthreads();

#pragma unroll 1
  for (int rl = 0; rl < nb; ++rl) {
    const float dkR = dk * rc[r0 + rl].R;
    const float dkRh = dkR * winvh;
    // (the unrolled fill_wtable measured slower here: it pushes this kernel
    // from 8 to 20 bytes of spills)
    for (int q = tid; q < ntab; q += kThreads) wtab[q] = filter_from_table(ktab[q], dkR, dkRh, invN3, wsh, wM);
    __syncthreads();
    // stage A: filter, radix 16 on the matrix cores, outputs to bufX
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int b = 4 * w + 2 * t + sel;
      float2 x[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int kx = b + 64 * (4 * g + j);
        const int fa = (kx <= kN / 2) ? kx : kN - kx;
        const float w0 = wtab[fa * kCF + c];
        if (!packed) {
          x[j] = cscale(m[4 * t + j], w0);
        } else {
          const float wn = wpk[fa];
          const float2 dn = DN[(long long)kx * kN + ky];
          x[j] = make_float2(w0 * m[4 * t + j].x - wn * dn.y, w0 * m[4 * t + j].y + wn * dn.x);
        }
      }
      const half8_t bop = pack_b(x);
      const float4_t z = {0.f, 0.f, 0.f, 0.f};
      const float4_t d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(gA0, bop, z, 0, 0, 0);
      const float4_t d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(gA1, bop, z, 0, 0, 0);
      // output q of butterfly b at 16 b + q
      store_d(bufX, d0, d1, 16 * b + g, 1, c);
    }
    __syncthreads();
    // stage B: k = w; butterflies b' = w + 16 (2t + sel), inputs at b' + 64 r
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int bp = w + 16 * (2 * t + sel);
      const half8_t bop = load_b(bufX, bp + 64 * 4 * g, 64, c);
      const float4_t z = {0.f, 0.f, 0.f, 0.f};
      const float4_t d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(gB0, bop, z, 0, 0, 0);
      const float4_t d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(gB1, bop, z, 0, 0, 0);
      // output q at (b' / 16) 256 + k + 16 q
      store_d(bufY, d0, d1, (bp >> 4) * 256 + w + 16 * g, 16, c);
    }
    __syncthreads();
    // stage C: radix 4, p = 256, twiddles w1024^(r b''), natural-order outputs
    __half2* dst = tp + (long long)rl * plane;
#pragma unroll
    for (int mm = 0; mm < 2; ++mm) {
      const int b2 = ic + 128 * mm;
      float2 u[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) u[r] = unpack_h2(bufY[xrow(b2 + 256 * r) * kCF + cl]);
#pragma unroll
      for (int r = 1; r < 4; ++r) u[r] = cmul(u[r], tw_full[256 + 4 * b2 + r]);
      DFT<4, +1>::run(u);
#pragma unroll
      for (int q = 0; q < 4; ++q) dst[(long long)(b2 + 256 * q) * kN * kC] = __float22half2_rn(u[q]);
    }
    // next radius: wtab was last read in stage A and bufX in stage B, both
    // before a barrier; bufY (read just above) is rewritten in stage B only
    // after the barriers that follow the table fill and stage A
  }
}

// ------------------------------------------------------ x pass, N = 2048
// The same pass for 2048 = 16 * 16 * 8 (FP16 intermediates): the two radix-16
// stages on the matrix cores with the same operand matrices (they depend on
// the stage's p only), the radix-8 stage (p = 256) on the vector ALUs.  The
// exchange buffers of 2048 rows leave room for 4 of the 8 kz columns of a
// spectrum tile (C = 8 at N = 2048) per block: MFMA column n = (column c = n
// & 3, butterfly sel = n >> 2).  Stage A: butterflies b = 8 w + 4 t + sel,
// inputs r = 4 g + j at kx = b + 128 r, outputs at 16 b + q.  Stage B: k =
// wave w, butterflies b' = w + 16 (4 t + sel), inputs at b' + 128 r, outputs at
// 256 (b' / 16) + w + 16 q.  Stage C: one radix-8 butterfly b2 = tid / 4 of
// column tid % 4 per thread, inputs at b2 + 256 r, twiddles w2048^(r b2) from
// the tile plan's stage table, outputs in natural order at b2 + 256 q.  |k| of
// the W table's entries is recomputed per radius (no |k| table: the W(x)
// table may take 56 KB of LDS here).
namespace x2048 {
constexpr int kN = 2048;
constexpr int kC = 8;            // Blk<2048>::C
constexpr int kCF = 4;           // columns per block
constexpr int kNSUB = kC / kCF;
constexpr int kNF = kN / 2 + 1;
constexpr int kThreads = 1024;
// Row swizzles (XOR of row bits 0-2, a bijection inside every 8 rows) that put
// the 32 lanes of every half-wave access on distinct banks: bufX is written
// at rows 16 b + q (lanes differ in q's bits 0-1 and b's bits 0-1 = rows' bits
// 4-5) and read at rows b' + 128 r (lanes differ in rows' bits 4-5 and 9);
// bufY is written at rows 256 j + k + 16 q (bits 4 and 8-9 vary) and read at
// rows b2 + 256 r (bits 0-2 vary).
__device__ __forceinline__ int xsw(int row) { return row ^ ((((row >> 4) & 3) << 1) ^ ((row >> 9) & 1)); }
__device__ __forceinline__ int ysw(int row) { return row ^ (((row >> 4) & 1) | (((row >> 8) & 3) << 1)); }
}  // namespace x2048

__global__ void __launch_bounds__(x2048::kThreads)
    k_fx_mfma2048(const float2* __restrict__ S, const float2* __restrict__ D0, const float2* __restrict__ DN,
                  __half2* __restrict__ Tbuf, float dk, float invN3, const RadiusConst* __restrict__ rc, int r0,
                  int nb, const float2* __restrict__ tw_plan, const float4* __restrict__ wcoef, int wM, float winvh,
                  const half8_t* __restrict__ mats) {
  constexpr int N2 = x2048::kN, C2 = x2048::kC, CF2 = x2048::kCF, NSUB2 = x2048::kNSUB, NF2 = x2048::kNF;
  constexpr int T2 = x2048::kThreads;
  using x2048::xsw;
  using x2048::ysw;
  constexpr long long plane = (long long)N2 * N2 * (N2 / 2);
  __shared__ uint32_t bufX[N2 * CF2];
  __shared__ uint32_t bufY[N2 * CF2];
  __shared__ float wtab[(CF2 + 1) * NF2];  // [|fx|][column], then the kz = N/2 column
  extern __shared__ float4 wsh[];           // W(x) interpolation table (wM intervals)
  float* wpk = wtab + CF2 * NF2;

  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int g = l >> 4, n = l & 15;
  const int c = n & 3, sel = n >> 2;

  const long long tile = xcd_tile(blockIdx.x, gridDim.x);
  const int sub = (int)(tile % NSUB2);
  const int ky = (int)((tile / NSUB2) % N2);
  const int kzb = (int)(tile / (NSUB2 * N2));
  const int kz0 = kzb * C2 + sub * CF2;
  const bool has_packed = (kzb == 0 && sub == 0);  // block-uniform
  const bool packed = has_packed && c == 0;        // this lane's column is kz = 0

  const half8_t gA0 = mats[mat_set_a(0) * 64 + l], gA1 = mats[mat_set_a(1) * 64 + l];
  const half8_t gB0 = mats[mat_set_b(w, 0) * 64 + l], gB1 = mats[mat_set_b(w, 1) * 64 + l];

  const float2* sp = S + ((long long)kzb * N2 + ky) * N2 * C2 + sub * CF2 + c;
  float2 m[8];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kx = 8 * w + 4 * t + sel + 128 * (4 * g + j);
      m[4 * t + j] = packed ? D0[(long long)kx * N2 + ky] : sp[(long long)kx * C2];
    }
  }
  const int fy2 = fsq(ky, N2);
  const int ntab = (has_packed ? CF2 + 1 : CF2) * NF2;
  for (int q = tid; q < wM; q += T2) wsh[q] = wcoef[q];
  const int cl = tid & 3, b2 = tid >> 2;  // stage C: column cl, butterfly b2 < 256
  __half2* tp = Tbuf + (long long)kzb * N2 * N2 * C2 + (long long)ky * C2 + sub * CF2 + cl;
  __syncthreads();

#pragma unroll 1
  for (int rl = 0; rl < nb; ++rl) {
    // opaque zero: the radius-invariant LDS and store addresses and the
    // stage-C twiddles are rebuilt per radius instead of held in registers
    int zo = 0;
    asm volatile("" : "+v"(zo));
    const float dkR = dk * rc[r0 + rl].R;
    const float dkRh = dkR * winvh;
    for (int q = tid; q < ntab; q += T2) {
      int f, kzq;
      if (q < CF2 * NF2) {
        f = q / CF2;
        kzq = kz0 + q % CF2;
      } else {
        f = q - CF2 * NF2;
        kzq = N2 / 2;
      }
      wtab[q] = filter_from_table(kmag_grid(f * f + fy2 + kzq * kzq), dkR, dkRh, invN3, wsh, wM);
    }
    __syncthreads();  // (1) W table complete; bufX's stage-B reads of the last radius done
    // stage A: filter, radix 16 on the matrix cores, outputs to bufX
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int b = 8 * w + 4 * t + sel;
      float2 x[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int kx = b + 128 * (4 * g + j);
        const int fa = (kx <= N2 / 2) ? kx : N2 - kx;
        const float w0 = wtab[fa * CF2 + c];
        if (!packed) {
          x[j] = cscale(m[4 * t + j], w0);
        } else {
          const float wn = wpk[fa];
          const float2 dn = DN[(long long)kx * N2 + ky];
          x[j] = make_float2(w0 * m[4 * t + j].x - wn * dn.y, w0 * m[4 * t + j].y + wn * dn.x);
        }
      }
      const half8_t bop = pack_b(x);
      const float4_t z = {0.f, 0.f, 0.f, 0.f};
      const float4_t d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(gA0, bop, z, 0, 0, 0);
      const float4_t d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(gA1, bop, z, 0, 0, 0);
      // outputs q = g, g + 4 (block 0), g + 8, g + 12 (block 1) at 16 b + q
      const int row = 16 * b + g + zo;
      bufX[xsw(row) * CF2 + c] = pack_h2(d0[0], d0[1]);
      bufX[xsw(row + 4) * CF2 + c] = pack_h2(d0[2], d0[3]);
      bufX[xsw(row + 8) * CF2 + c] = pack_h2(d1[0], d1[1]);
      bufX[xsw(row + 12) * CF2 + c] = pack_h2(d1[2], d1[3]);
    }
    __syncthreads();  // (2) stage A's outputs complete
    // stage B: k = w; butterflies b' = w + 16 (4 t + sel), inputs at b' + 128 r
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int bp = w + 16 * (4 * t + sel);
      union {
        uint32_t u[4];
        half8_t h;
      } cv;
#pragma unroll
      for (int j = 0; j < 4; ++j) cv.u[j] = bufX[xsw(bp + zo + 128 * (4 * g + j)) * CF2 + c];
      const float4_t z = {0.f, 0.f, 0.f, 0.f};
      const float4_t d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(gB0, cv.h, z, 0, 0, 0);
      const float4_t d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(gB1, cv.h, z, 0, 0, 0);
      // output q at 256 (b' / 16) + w + 16 q
      const int row = (bp >> 4) * 256 + w + 16 * g + zo;
      bufY[ysw(row) * CF2 + c] = pack_h2(d0[0], d0[1]);
      bufY[ysw(row + 64) * CF2 + c] = pack_h2(d0[2], d0[3]);
      bufY[ysw(row + 128) * CF2 + c] = pack_h2(d1[0], d1[1]);
      bufY[ysw(row + 192) * CF2 + c] = pack_h2(d1[2], d1[3]);
    }
    __syncthreads();  // (3) stage B's outputs complete
    // stage C: radix 8, p = 256, natural-order outputs
    __half2* dst = tp + (long long)rl * plane;
    float2 u[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) u[r] = unpack_h2(bufY[ysw(b2 + zo + 256 * r) * CF2 + cl]);
#pragma unroll
    for (int r = 1; r < 8; ++r) u[r] = cmul(u[r], tw_plan[256 + 8 * b2 + r + zo]);
    DFT<8, +1>::run(u);
#pragma unroll
    for (int q = 0; q < 8; ++q) dst[(long long)(b2 + zo + 256 * q) * N2 * C2] = __float22half2_rn(u[q]);
    // next radius: wtab was last read in stage A and bufX in stage B, both
    // before a barrier; bufY (read just above) is rewritten in stage B only
    // after the barriers that follow the table fill and stage A
  }
}

// ---------------------------------------------------------------- z pass
// Last inverse pass (z, c2r) of the mixed path at N = 1024 with the barrier
// fused, as k_fz (reference delta2zreion_d2z.C:92-146,206-216), for blocks of
// 8 consecutive z rows (512 B contiguous per kz block in the T layout).  Per
// radius: the raw FP16 rows are staged in LDS; stage A forms the complex
// sequence Z[k] of the half-length c2r from X[k] and X[L-k] (packed FP16) and
// runs radix 16 on the matrix cores; stage B (radix 16, p = 16) groups the 16
// butterflies {k, k+16} x 8 rows that share k = b mod 16; stage C (radix 2,
// p = 256, FP32) leaves row w in wave w, which evaluates d2z() in FP32 (one FMA
// per cell), keeps the maximum over the batch in registers and the per-radius
// sums of squares.
constexpr int kLz = kN / 2;   // complex points per z row
constexpr int kRowsZ = 8;     // rows per block
// Padded layouts (bank model: conflict-free except stage B's writes, 2-way);
// every access is a per-lane base plus a compile-time offset, except the
// mirrored reads X[L-k]
constexpr int kXStride = 524, kY1Stride = 580, kY2Stride = 516;
__device__ __forceinline__ int zx_at(int row, int pos) {  // raw rows X
  return row * kXStride + pos + 2 * (pos >> 7);
}
__device__ __forceinline__ int zy1_at(int row, int pos) {  // stage A output
  return row * kY1Stride + pos + 2 * (pos >> 4) + (pos >> 7);
}
__device__ __forceinline__ int zy2_at(int row, int pos) {  // stage B output (aliases X)
  return row * kY2Stride + pos + (pos >> 7) + (pos >> 8);
}

template <bool RB>
__global__ void __launch_bounds__(512)
    k_fz_mfma(const __half2* __restrict__ Tbuf, float* __restrict__ zre, float* __restrict__ rbub,
              const RadiusConst* __restrict__ rc, int r0, int nb, double* __restrict__ partial, float iscale,
              const float2* __restrict__ tw_full, const half8_t* __restrict__ mats) {
  constexpr long long nrows = (long long)kN * kN;
  constexpr long long plane = nrows * kLz;
  __shared__ uint32_t bufX[kRowsZ * kXStride];  // also holds stage B's output (8 x 516)
  __shared__ uint32_t bufY[kRowsZ * kY1Stride];
  static_assert(kY2Stride <= kXStride, "stage B output must fit in the raw buffer");
  __shared__ uint2 s_pq[kLz];        // glue constants P, Q (FP16 pairs), k < N/2
  __shared__ float2 s_tw2[kLz / 2];  // exp(2 pi i b / (N/2)) (stage C)
  __shared__ double red[kRowsZ][kFusedMaxRadii];
  __shared__ BarrierF32 s_bk[kFusedMaxRadii];  // barrier constants of the batch's radii
  // operand matrices in the A layout, per lane: stage A [h][lane], stage B
  // [wave][t][h][lane] (k = 2 wave + t); read per radius rather than held in
  // registers across radii
  __shared__ half8_t s_gA[2][64];
  __shared__ half8_t s_gB[kRowsZ][2][2][64];
  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int g = l >> 4, n = l & 15;
  const int rowA = n & 7, sel = n >> 3;  // MFMA column n = (row, butterfly sel)
  if (tid < kRowsZ * kFusedMaxRadii) red[tid / kFusedMaxRadii][tid % kFusedMaxRadii] = 0.0;
  for (int q = tid; q < kLz; q += 512) {
    s_pq[q] = c2r_glue_pq(q, tw_full[q]);
  }
  for (int q = tid; q < kLz / 2; q += 512) s_tw2[q] = tw_full[2 * q];
  if (tid < nb) s_bk[tid] = barrier_f32(rc[r0 + tid].A, rc[r0 + tid].B, iscale);
  if (w < 2) s_gA[w][l] = mats[mat_set_a(w) * 64 + l];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    s_gB[w][t][0][l] = mats[mat_set_b(2 * w + t, 0) * 64 + l];
    s_gB[w][t][1][l] = mats[mat_set_b(2 * w + t, 1) * 64 + l];
  }
  const double sc = (double)iscale, sc2 = sc * sc;
  // raw-load mapping: each half wave one row: kz = 16 kzb + (t & 15) with
  // kzb = ((t >> 4) & 1) + 2 (t >> 8) + 4 e, row (t >> 5) & 7
  const int lrow = (tid >> 5) & 7, lkz = tid & 15, lkzb = ((tid >> 4) & 1) + 2 * (tid >> 8);
  const float4_t z4 = {0.f, 0.f, 0.f, 0.f};
  // the raw rows of (group, radius) are loaded into registers one step ahead
  // and stored to LDS at the start of the step, so the loads overlap the
  // previous step's transforms
  uint32_t pf[8];
  const uint32_t lane_off = ((uint32_t)lkzb * (uint32_t)nrows + (uint32_t)lrow) * 16u + (uint32_t)lkz;
  auto fetch = [&](long long gbase, int q) {
    const uint32_t* tp = reinterpret_cast<const uint32_t*>(Tbuf + (long long)q * plane + gbase * 16);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      // uniform part (kz block 4e) on the scalar base, one per-lane offset
      const uint32_t* te = tp + (size_t)(4 * e) * nrows * 16;
      pf[e] = te[lane_off];
    }
  };
  const long long stride = (long long)gridDim.x * kRowsZ;
  if ((long long)blockIdx.x * kRowsZ < nrows) fetch((long long)blockIdx.x * kRowsZ, 0);
  __syncthreads();
#pragma unroll 1
  for (long long base = (long long)blockIdx.x * kRowsZ; base < nrows; base += stride) {
    float2 best[8];
    float bR0[8], bR1[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      best[e] = make_float2(0.0f, 0.0f);  // zreion >= 0: see zmax_nonneg
      bR0[e] = 0.f;
      bR1[e] = 0.f;
    }
#pragma unroll 1
    for (int q = 0; q < nb; ++q) {
      int zo = 0;  // opaque zero: the matrix reads stay inside the radius loop
      asm volatile("" : "+s"(zo));
#pragma unroll
      for (int e = 0; e < 8; ++e) bufX[zx_at(lrow, 16 * (lkzb + 4 * e) + lkz)] = pf[e];
      __syncthreads();
      // stage A: butterflies b = 4w + 2t + sel of row rowA, inputs Z[b + 32 r]
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int b = 4 * w + 2 * t + sel;
        union {
          uint32_t u[4];
          half8_t h;
        } zz;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = b + 32 * (4 * g + j);
          const uint2 pq = s_pq[k];
          zz.u[j] = c2r_glue_h2(bufX[zx_at(rowA, k)], bufX[zx_at(rowA, (kLz - k) & (kLz - 1))], pq.x, pq.y);
        }
        const half8_t bop = zz.h;
        const float4_t d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(s_gA[0][l + zo], bop, z4, 0, 0, 0);
        const float4_t d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(s_gA[1][l + zo], bop, z4, 0, 0, 0);
        const int p0 = 16 * b + g;
        bufY[zy1_at(rowA, p0)] = pack_h2(d0[0], d0[1]);
        bufY[zy1_at(rowA, p0 + 4)] = pack_h2(d0[2], d0[3]);
        bufY[zy1_at(rowA, p0 + 8)] = pack_h2(d1[0], d1[1]);
        bufY[zy1_at(rowA, p0 + 12)] = pack_h2(d1[2], d1[3]);
      }
      __syncthreads();
      if (q + 1 < nb)
        fetch(base, q + 1);
      else if (base + stride < nrows)
        fetch(base + stride, 0);
      // stage B: k = 2w + t, butterflies b' = k + 16 sel of row rowA, inputs
      // b' + 32 r; outputs to the raw buffer, free since stage A's barrier
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int kk = 2 * w + t, bp = kk + 16 * sel;
        union {
          uint32_t u[4];
          half8_t h;
        } cv;
#pragma unroll
        for (int j = 0; j < 4; ++j) cv.u[j] = bufY[zy1_at(rowA, bp + 32 * (4 * g + j))];
        const float4_t d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(s_gB[w][t][0][l + zo], cv.h, z4, 0, 0, 0);
        const float4_t d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(s_gB[w][t][1][l + zo], cv.h, z4, 0, 0, 0);
        const int p0 = sel * 256 + kk + 16 * g;
        bufX[zy2_at(rowA, p0)] = pack_h2(d0[0], d0[1]);
        bufX[zy2_at(rowA, p0 + 64)] = pack_h2(d0[2], d0[3]);
        bufX[zy2_at(rowA, p0 + 128)] = pack_h2(d1[0], d1[1]);
        bufX[zy2_at(rowA, p0 + 192)] = pack_h2(d1[2], d1[3]);
      }
      __syncthreads();
      // stage C: row w, butterflies b'' = l + 64 m; outputs n = b'' and b'' + 256
      float2 v[8];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int b2 = l + 64 * m;
        const float2 u0 = unpack_h2(bufX[zy2_at(w, b2)]);
        const float2 u1 = cmul(unpack_h2(bufX[zy2_at(w, b2 + 256)]), s_tw2[b2]);
        v[m] = cadd(u0, u1);
        v[m + 4] = csub(u0, u1);
      }
      lds_hold(v);      // the reads of bufX stay above the barrier
      __syncthreads();  // the raw buffer is reloaded next radius
      const BarrierF32 bk = s_bk[q];
      const float R = rc[r0 + q].R;
      float acc = 0.0f;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float z0 = barrier_z_f32(v[e].x, bk);
        const float z1 = barrier_z_f32(v[e].y, bk);
        acc = fmaf(v[e].x, v[e].x, fmaf(v[e].y, v[e].y, acc));
        if constexpr (RB) {
          if (z0 > best[e].x) {
            best[e].x = z0;
            bR0[e] = R;
          }
          if (z1 > best[e].y) {
            best[e].y = z1;
            bR1[e] = R;
          }
        } else {
          best[e].x = zmax_nonneg(best[e].x, z0);
          best[e].y = zmax_nonneg(best[e].y, z1);
        }
      }
      const double wsum = (double)wave_sum_f(acc) * sc2;
      if (l == 0) red[w][q] += wsum;
    }
    // zreion (and rbubble) of row base + w, element pair n = l + 64 m (+ 256)
    float2* zp = reinterpret_cast<float2*>(zre + (base + w) * kN);
    float2* bp2 = RB ? reinterpret_cast<float2*>(rbub + (base + w) * kN) : nullptr;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int nn = l + 64 * (e & 3) + 256 * (e >> 2);
      const float2 old = zp[nn];
      const bool u0 = best[e].x > old.x, u1 = best[e].y > old.y;
      if (u0 || u1) {
        zp[nn] = make_float2(u0 ? best[e].x : old.x, u1 ? best[e].y : old.y);
        if (RB) {
          const float2 ob = bp2[nn];
          bp2[nn] = make_float2(u0 ? bR0[e] : ob.x, u1 ? bR1[e] : ob.y);
        }
      }
    }
  }
  __syncthreads();
  if (tid < nb) {
    double s = 0.0;
    for (int k = 0; k < kRowsZ; ++k) s += red[k][tid];
    partial[(long long)(r0 + tid) * gridDim.x + blockIdx.x] = s;
  }
}

}  // namespace

bool fz_mfma_supported(int N, bool half) { return half && N == kN; }

int fz_mfma_blocks(bool rb) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = rb ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fz_mfma<true>, 512, 0)
           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fz_mfma<false>, 512, 0);
  long long g = (e == hipSuccess && cus > 0 && per_cu > 0) ? (long long)cus * per_cu : 1024;
  if (g > kFzMaxBlocks) g = kFzMaxBlocks;
  const long long groups = (long long)kN * kN / kRowsZ;
  return (int)(g < groups ? g : groups);
}

hipError_t launch_fz_mfma(int N, const void* Tbuf, float* zre, float* rbub, const RadiusConst* rc, int r0, int nb,
                          double* partial, int nblk, float iscale, const float2* tw_full, const void* mats,
                          hipStream_t s) {
  if (N != kN || nb < 1 || nb > kFusedMaxRadii || nblk < 1 || nblk > kFzMaxBlocks || !mats)
    return hipErrorInvalidValue;
  const __half2* T = static_cast<const __half2*>(Tbuf);
  const half8_t* M = static_cast<const half8_t*>(mats);
  if (rbub)
    hipLaunchKernelGGL(k_fz_mfma<true>, dim3(nblk), dim3(512), 0, s, T, zre, rbub, rc, r0, nb, partial, iscale,
                       tw_full, M);
  else
    hipLaunchKernelGGL(k_fz_mfma<false>, dim3(nblk), dim3(512), 0, s, T, zre, rbub, rc, r0, nb, partial, iscale,
                       tw_full, M);
  return hipGetLastError();
}

bool fx_mfma_supported(int N, bool half) { return half && (N == kN || N == x2048::kN); }

hipError_t launch_fx_mfma(int N, const float2* S, const float2* D0, const float2* DN, void* Tbuf, float dk,
                          float invN3, const RadiusConst* rc, int r0, int nb, const float2* tw_full_plan,
                          const WTable& wt, const void* mats, hipStream_t s) {
  if (N == x2048::kN) {
    if (nb < 1 || nb > kFxMaxRadii || wt.M < 0 || wt.M > kWTableMax || !mats) return hipErrorInvalidValue;
    const long long nblk = (long long)x2048::kN * ((x2048::kN / 2) / x2048::kC) * x2048::kNSUB;
    hipLaunchKernelGGL(k_fx_mfma2048, dim3((unsigned)nblk), dim3(x2048::kThreads), sizeof(float4) * wt.M, s, S, D0,
                       DN, static_cast<__half2*>(Tbuf), dk, invN3, rc, r0, nb, tw_full_plan, wt.coef, wt.M, wt.invh,
                       static_cast<const half8_t*>(mats));
    return hipGetLastError();
  }
  if (N != kN || nb < 1 || nb > kFxMaxRadii || wt.M < 0 || wt.M > kWTableMax || !mats) return hipErrorInvalidValue;
  const long long nblk = (long long)kN * ((kN / 2) / kC) * kNSUB;
  hipLaunchKernelGGL(k_fx_mfma, dim3((unsigned)nblk), dim3(kThreads), sizeof(float4) * wt.M, s, S, D0, DN,
                     static_cast<__half2*>(Tbuf), dk, invN3, rc, r0, nb, tw_full_plan, wt.coef, wt.M, wt.invh,
                     static_cast<const half8_t*>(mats));
  return hipGetLastError();
}

size_t mfma_matrix_bytes(int N, bool half) {
  return (fx_mfma_supported(N, half) || fz_mfma_supported(N, half)) ? sizeof(half8_t) * kMatSets * 64 : 0;
}

hipError_t launch_mfma_matrices(void* out, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_mats, dim3(kMatSets), dim3(64), 0, s, static_cast<half8_t*>(out));
  return hipGetLastError();
}

}  // namespace prc
