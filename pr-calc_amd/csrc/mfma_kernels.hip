// Matrix-core (MFMA) variant of the fused x pass for the mixed path at
// N = 1024 (FP16 intermediates): the two radix-16 stages of the 1024-point
// inverse transform (1024 = 16 * 16 * 4) run as 32 x 32 real matrix products
// on v_mfma_f32_16x16x32_f16 (FP16 operands, FP32 accumulation), the radix-4
// stage on the vector ALUs.  Reference: delta2zreion_d2z.C:58-92 (filter and
// c2r), with the same W(kR) table and the same output layout as k_fx.
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

__device__ __forceinline__ uint32_t pack_h2(float re, float im) {
  const __half2 h = __floats2half2_rn(re, im);
  return *reinterpret_cast<const uint32_t*>(&h);
}
__device__ __forceinline__ float2 unpack_h2(uint32_t u) {
  __half2 h;
  *reinterpret_cast<uint32_t*>(&h) = u;
  return __half22float2(h);
}

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

__global__ void __launch_bounds__(kThreads)
    k_fx_mfma(const float2* __restrict__ S, const float2* __restrict__ D0, const float2* __restrict__ DN,
              __half2* __restrict__ Tbuf, float dk, float invN3, const RadiusConst* __restrict__ rc, int r0, int nb,
              const float2* __restrict__ tw_full, const float4* __restrict__ wcoef, int wM, float winvh) {
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
  const half8_t gA0 = stage_matrix(0, 1, 0, l), gA1 = stage_matrix(0, 1, 1, l);
  const half8_t gB0 = stage_matrix(w, 16, 0, l), gB1 = stage_matrix(w, 16, 1, l);

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

}  // namespace

bool fx_mfma_supported(int N, bool half) { return half && N == kN; }

hipError_t launch_fx_mfma(int N, const float2* S, const float2* D0, const float2* DN, void* Tbuf, float dk,
                          float invN3, const RadiusConst* rc, int r0, int nb, const float2* tw_full_plan,
                          const WTable& wt, hipStream_t s) {
  if (N != kN || nb < 1 || nb > kFusedMaxRadii || wt.M < 0 || wt.M > kWTableMax) return hipErrorInvalidValue;
  const long long nblk = (long long)kN * ((kN / 2) / kC) * kNSUB;
  hipLaunchKernelGGL(k_fx_mfma, dim3((unsigned)nblk), dim3(kThreads), sizeof(float4) * wt.M, s, S, D0, DN,
                     static_cast<__half2*>(Tbuf), dk, invN3, rc, r0, nb, tw_full_plan, wt.coef, wt.M, wt.invh);
  return hipGetLastError();
}

}  // namespace prc
