// Compensated matrix-core passes of the mixed path at N = 1024: FP32
// intermediates, radix-16 DFT stages on v_mfma_f32_16x16x32_f16 with split
// hi/lo operands and FP32 accumulation.
//
// Every FP32 operand v is split into two FP16 values, v_hi = fp16(v) and
// v_lo = fp16((v - v_hi) 2^11), so that v = v_hi + 2^-11 v_lo to 22 bits
// (v_lo is scaled so it stays a normal FP16 number wherever v does).  A product
// of a matrix A and data X is then
//     A X = A_hi X_hi + 2^-11 (A_hi X_lo + A_lo X_hi)   (+ 2^-22 A_lo X_lo, dropped)
// three MFMAs, the two 2^-11 terms accumulated into their own FP32 registers
// and added with one FMA per output.  The matrices are split once at set-up
// from FP64 values.  Products of FP16 values are exact in FP32, so the error is
// that of the 22-bit operands and the FP32 accumulation: FP32-grade, not FP16.
//
// Reference: the c2r of delta2zreion_d2z.C:92 with the filter (:58-89) fused
// into the x pass and the barrier and running maximum (:106-146, :206-216) into
// the z pass, as the FP32 fused kernels.
//
// Stockham formulation (as fft_core.h): stage (p, R) takes butterfly b's
// inputs x[b + r L/R], multiplies input r by w^(r (b mod p)) with w =
// exp(+2 pi i / (p R)), applies the R-point DFT and writes output q to
// y[(b/p) p R + (b mod p) + q p].  MFMA operand layout (16x16x32, lane l):
// A[l % 16][8 (l/16) + j], B[8 (l/16) + j][l % 16], D[4 (l/16) + i][l % 16].
// A radix-16 DFT of 16 butterflies is the 32 x 32 real matrix (re/im
// interleaved) times a 32 x 16 operand: B row K = 8 g + 2 j + c holds input
// r = 4 g + j, component c, of butterfly column n; A row M = 16 h + 4 g' + 2 s + c'
// gives output q = g' + 4 s + 8 h.  Lane l = 16 g + n supplies 4 complex inputs
// of butterfly n and receives outputs q = g, g + 4 (block h = 0) and g + 8,
// g + 12 (h = 1).
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

constexpr int kN = 1024;         // grid size and x-transform length
constexpr int kC = 16;           // blocked-layout width (Blk<1024>::C)
constexpr int kCF = 8;           // columns (kz) per x block
constexpr int kNSUB = kC / kCF;  // column sub-tiles per kz block
constexpr int kNF = kN / 2 + 1;  // distinct |fx|
constexpr float kLoScale = 2048.0f;
constexpr float kLoInv = 1.0f / 2048.0f;

// ------------------------------------------------------ split operands
struct SplitOp {
  half8_t hi, lo;
};

// one 16-row block h of the radix-16 stage matrix G_k (stage p), lane l's A
// operand, split from FP64: angle 2 pi (q r / 16 + r k / (16 p))
__device__ SplitOp split_stage_matrix(int k, int p, int h, int l) {
  const int m = l & 15, gp = l >> 4;
  const int g = m >> 2, s = (m >> 1) & 1, c = m & 1;
  const int q = g + 4 * s + 8 * h;
  SplitOp o;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int r = 4 * gp + (j >> 1), cp = j & 1;
    const int theta = (p * q * r + r * k) % (16 * p);
    double sn, cs;
    sincospi((double)theta / (double)(8 * p), &sn, &cs);
    const double v = (c == 0) ? (cp == 0 ? cs : -sn) : (cp == 0 ? sn : cs);
    const _Float16 hv = (_Float16)v;
    o.hi[j] = hv;
    o.lo[j] = (_Float16)((v - (double)hv) * 2048.0);
  }
  return o;
}

// operand sets: 0, 1 = stage p = 1 (the plain 16-point DFT) blocks h = 0, 1;
// 2 + 2 k + h = stage p = 16, twiddle index k, block h; stored
// [set][part][lane] with part 0 = hi, 1 = lo 2^11
constexpr int kSplitSets = 2 + 2 * 16;
constexpr int kSplitParts = 2;

__global__ void __launch_bounds__(64) k_split_mats(half8_t* __restrict__ out) {
  const int set = blockIdx.x, l = threadIdx.x;
  const SplitOp o = set < 2 ? split_stage_matrix(0, 1, set, l) : split_stage_matrix((set - 2) >> 1, 16, (set - 2) & 1, l);
  out[(set * kSplitParts + 0) * 64 + l] = o.hi;
  out[(set * kSplitParts + 1) * 64 + l] = o.lo;
}

__device__ __forceinline__ SplitOp load_mat(const half8_t* __restrict__ mats, int set, int l) {
  return SplitOp{mats[(set * kSplitParts + 0) * 64 + l], mats[(set * kSplitParts + 1) * 64 + l]};
}

// data split of 4 complex values into the B operand pair.  The low part
// 2^11 (x - hi) is formed by one mixed-precision FMA from the FP16 high part,
// fma(hi, -2^11, 2^11 x), and rounded to FP16 in the same instruction
// (v_fma_mixlo/mixhi_f16): 2^11 (x - hi) is exact in FP32, so the FMA's single
// rounding gives the same value as the subtract and scale, with 52 fewer vector
// instructions per radius in the x pass (the pass's time is set by its stores
// and did not change measurably).
__device__ __forceinline__ SplitOp split_b(const float2 (&x)[4]) {
  SplitOp o;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const _Float16 hr = (_Float16)x[j].x, hi = (_Float16)x[j].y;
    o.hi[2 * j] = hr;
    o.hi[2 * j + 1] = hi;
    o.lo[2 * j] = (_Float16)__builtin_fmaf((float)hr, -kLoScale, x[j].x * kLoScale);
    o.lo[2 * j + 1] = (_Float16)__builtin_fmaf((float)hi, -kLoScale, x[j].y * kLoScale);
  }
  return o;
}

// A X with the split operands: three MFMAs, FP32 accumulation
__device__ __forceinline__ float4_t mfma_split(const SplitOp& a, const SplitOp& b) {
  const float4_t z = {0.f, 0.f, 0.f, 0.f};
  float4_t c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.hi, b.lo, z, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.lo, b.hi, c, 0, 0, 0);
  const float4_t d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.hi, b.hi, z, 0, 0, 0);
  float4_t r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r[i] = fmaf(c[i], kLoInv, d[i]);
  return r;
}

// ---------------------------------------------------------------- x pass
// Filter + x inverse for nb radii.  A block owns 8 of the 16 kz columns of one
// spectrum tile (ky, kz block) and all 1024 kx; 16 waves.  Per radius: the W
// table of the block's columns (LDS), stage A (p = 1) straight from the
// registers holding the spectrum, one LDS exchange (FP32), stage B (p = 16,
// k = wave index, so each wave keeps G_k in registers), then the radix-4 stage
// (p = 256) in registers after one exchange with lane l ^ 8 and the stores of
// the FP32 intermediate T_r (T layout, natural order along x).
//
// Operand range: the block scales its spectrum by a power of two bs chosen
// once per launch so that |filtered value| <= 2^10; the radix-16 stage grows
// values by at most 16 sqrt 2, so the second stage's FP16 parts stay below
// 2^15.  The outputs are multiplied by 1/bs (exact) before they are stored.
// The two correction products keep their own accumulator and join the main
// product in one FP32 FMA.  Accumulating all three products in one MFMA
// accumulator raised the largest error of a 1024^3 smoothed field from 2.7e-6
// to 5.3e-5 - 5.8e-5 of its rms, whether the low parts were left unscaled or
// the matrices (x 2^5) and the data (<= 2^5) were scaled so that every part is
// a normal FP16 number: the matrix cores' sum of the large products and the
// small accumulator input is not FP32-exact.
//
// LDS exchange rows hold 8 columns of float2 (64 B); row `row` is stored at
// row ^ bit4(row), which puts the two halves of every quarter wave (butterfly
// parity sel = bit 4 of the row in both access patterns) on disjoint banks.
__device__ __forceinline__ int xrow_s(int row) { return row ^ ((row >> 4) & 1); }

constexpr int kFxThreads = 1024;

__global__ void __launch_bounds__(kFxThreads)
    k_fx_split(const float2* __restrict__ S, const float2* __restrict__ D0, const float2* __restrict__ DN,
               float2* __restrict__ Tbuf, float dk, float invN3, const RadiusConst* __restrict__ rc, int r0, int nb,
               const float2* __restrict__ tw_full, const float4* __restrict__ wcoef, int wM, float winvh,
               const half8_t* __restrict__ mats) {
  constexpr long long Nh = kN / 2;
  constexpr long long plane = (long long)kN * kN * Nh;
  __shared__ float2 bufX[kN * kCF];
  __shared__ float wtab[(kCF + 1) * kNF];  // [|fx|][column], then the kz = N/2 column
  __shared__ float ktab[(kCF + 1) * kNF];  // |k|/dk of the same entries
  __shared__ float s_max[kFxThreads / 64];
  __shared__ half8_t s_gA[2][2][64];  // the 16-point DFT (stage A), [block h][hi/lo][lane]
  __shared__ float2 s_twc[3 * 256];   // exp(2 pi i m / 1024), m < 768 (stage C)
  extern __shared__ float4 wsh[];     // W(x) interpolation table (wM intervals)
  float* wpk = wtab + kCF * kNF;

  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int g = l >> 4, n = l & 15;
  const int c = n & 7, sel = n >> 3;  // MFMA column n = (column c, butterfly parity sel)

  const long long tile = line_tile(blockIdx.x, gridDim.x, line_group(kNSUB, kC * (int)sizeof(float2)));
  const int sub = (int)(tile % kNSUB);
  const int ky = (int)((tile / kNSUB) % kN);
  const int kzb = (int)(tile / (kNSUB * kN));
  const int kz0 = kzb * kC + sub * kCF;
  const bool has_packed = (kzb == 0 && sub == 0);  // block-uniform
  const bool packed = has_packed && c == 0;        // this lane's column is kz = 0

  // stage B's matrices G_k (k = w) stay in registers; stage A's in LDS
  const SplitOp gB0 = load_mat(mats, 2 + 2 * w, l), gB1 = load_mat(mats, 3 + 2 * w, l);
  if (w < 4) s_gA[w >> 1][w & 1][l] = mats[((w >> 1) * kSplitParts + (w & 1)) * 64 + l];
  for (int q = tid; q < 3 * 256; q += kFxThreads) s_twc[q] = tw_full[q];

  // the lane's 8 spectrum values: stage-A butterflies b = 4w + 2t + sel,
  // inputs r = 4g + j at kx = b + 64 r
  const float2* sp = S + ((long long)kzb * kN + ky) * kN * kC + sub * kCF + c;
  float2 m[8];
  float mx = 0.0f;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kx = 4 * w + 2 * t + sel + 64 * (4 * g + j);
      m[4 * t + j] = packed ? D0[(long long)kx * kN + ky] : sp[(long long)kx * kC];
      mx = fmaxf(mx, fmaxf(fabsf(m[4 * t + j].x), fabsf(m[4 * t + j].y)));
      if (packed) {
        const float2 dn = DN[(long long)kx * kN + ky];
        mx = fmaxf(mx, fmaxf(fabsf(dn.x), fabsf(dn.y)));
      }
    }
  }
  // block scale: |W| <= 1, so |filtered| <= 2 mx invN3 (the packed column sums
  // two terms); bs puts that below 2^10
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  if (l == 0) s_max[w] = mx;
  const int fy2 = fsq(ky, kN);
  const int ntab = (has_packed ? kCF + 1 : kCF) * kNF;
  for (int q = tid; q < ntab; q += kFxThreads) {
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
  for (int q = tid; q < wM; q += kFxThreads) wsh[q] = wcoef[q];
  const int qbase = sel ? 2 : 0;
  __syncthreads();
  float bmax = 0.0f;
#pragma unroll
  for (int k = 0; k < kFxThreads / 64; ++k) bmax = fmaxf(bmax, s_max[k]);
  int e2 = 0;
  (void)frexpf(2.0f * bmax * invN3, &e2);  // 2 bmax invN3 < 2^e2
  e2 = bmax > 0.0f ? min(max(e2, -100), 100) : 10;
  const float bs = ldexpf(1.0f, 10 - e2), ibs = ldexpf(1.0f, e2 - 10);
  const float wscale = invN3 * bs;

  float2* tp = Tbuf + (long long)kzb * kN * kN * kC + (long long)ky * kC + sub * kCF + c;

#pragma unroll 1
  for (int rl = 0; rl < nb; ++rl) {
    int zo = 0;  // opaque zero: the LDS operand reads stay inside the radius loop
    asm volatile("" : "+v"(zo));
    const float dkR = dk * rc[r0 + rl].R;
    const float dkRh = dkR * winvh;
    fill_wtable<kFxThreads, ((kCF + 1) * kNF + kFxThreads - 1) / kFxThreads>(wtab, ktab, ntab, dkR, dkRh, wscale, wsh,
                                                                              wM, tid);
    __syncthreads();  // (1) W table complete; the previous radius's stage-B reads of bufX are done
    // stage A: filter, split, radix 16 on the matrix cores, outputs to bufX
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
          const float2 dn = DN[(long long)(kx + zo) * kN + ky];  // address rebuilt per radius
          x[j] = make_float2(w0 * m[4 * t + j].x - wn * dn.y, w0 * m[4 * t + j].y + wn * dn.x);
        }
      }
      const SplitOp bop = split_b(x);
      const float4_t d0 = mfma_split(SplitOp{s_gA[0][0][l + zo], s_gA[0][1][l + zo]}, bop);
      const float4_t d1 = mfma_split(SplitOp{s_gA[1][0][l + zo], s_gA[1][1][l + zo]}, bop);
      // output q of butterfly b at row 16 b + q
      const int row = 16 * b + g;
      bufX[xrow_s(row) * kCF + c] = make_float2(d0[0], d0[1]);
      bufX[xrow_s(row + 4) * kCF + c] = make_float2(d0[2], d0[3]);
      bufX[xrow_s(row + 8) * kCF + c] = make_float2(d1[0], d1[1]);
      bufX[xrow_s(row + 12) * kCF + c] = make_float2(d1[2], d1[3]);
    }
    __syncthreads();  // (2) stage A's outputs complete; its W table reads done
    // stage B: k = w; butterflies b' = w + 16 (2t + sel), inputs at b' + 64 r;
    // outputs q = g + 4 qi of butterfly t at (2t + sel) 256 + w + 16 q stay in
    // registers (named values, not an array: no stack copy)
    float4_t yb[2][2];  // [t][block h]: outputs qi = 2h (d[0], d[1]) and 2h + 1 (d[2], d[3])
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int bp = w + 16 * (2 * t + sel);
      float2 x[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) x[j] = bufX[xrow_s(bp + 64 * (4 * g + j)) * kCF + c];
      const SplitOp bop = split_b(x);
      yb[t][0] = mfma_split(gB0, bop);
      yb[t][1] = mfma_split(gB1, bop);
    }
    // stage C: radix 4, p = 256.  Butterfly b2 = w + 16 (g + 4 qi) takes
    // inputs r = 2t + sel' from this lane (sel' = sel) and lane l ^ 8; this
    // lane computes qi = qbase + u (u < 2): block h = sel of its own outputs,
    // and sends the partner block h = 1 - sel.  The exchange is a DPP row
    // rotation by 8 (lane l <-> l ^ 8 within each row of 16).
    float4_t mine[2], recv[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      float4_t send;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        mine[t][i] = sel ? yb[t][1][i] : yb[t][0][i];
        send[i] = sel ? yb[t][0][i] : yb[t][1][i];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
        recv[t][i] = __int_as_float(
            __builtin_amdgcn_update_dpp(0, __float_as_int(send[i]), 0x128 /* row_ror:8 */, 0xf, 0xf, false));
    }
    float2* dst = tp + (long long)rl * plane;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int qi = qbase + u;
      const int b2 = w + 16 * (g + 4 * qi);
      const float2 y0 = make_float2(mine[0][2 * u], mine[0][2 * u + 1]);
      const float2 y1 = make_float2(mine[1][2 * u], mine[1][2 * u + 1]);
      const float2 p0 = make_float2(recv[0][2 * u], recv[0][2 * u + 1]);
      const float2 p1 = make_float2(recv[1][2 * u], recv[1][2 * u + 1]);
      float2 v[4];
      v[0] = sel ? p0 : y0;
      v[1] = sel ? y0 : p0;
      v[2] = sel ? p1 : y1;
      v[3] = sel ? y1 : p1;
#pragma unroll
      for (int r = 1; r < 4; ++r) v[r] = cmul(v[r], s_twc[r * b2 + zo]);
      DFT<4, +1>::run(v);
      // Stores of 16 B: lanes c and c ^ 1 (same rows) swap one value, then
      // the even lane writes columns (c, c + 1) of row q = 2p and the odd lane
      // columns (c - 1, c) of row 2p + 1.  Same bits as a lane per column;
      // at 1024^3 x 96 the x pass measured 126.5-129.2 against 130.3-140.9 ms
      // per step (the y pass after it 2-3 ms slower, the step faster).  The
      // offsets are rebuilt from the opaque zero each radius (not hoisted out
      // of the loop and held in registers).
      const bool odd = (c & 1) != 0;
      float2* const dste = dst - (odd ? 1 : 0);
#pragma unroll
      for (int pp = 0; pp < 2; ++pp) {
        const float2 a = cscale(v[2 * pp], ibs), bq = cscale(v[2 * pp + 1], ibs);
        const float2 snd = odd ? a : bq;
        const float2 rcv = make_float2(
            __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(snd.x), 0xB1 /* quad_perm [1,0,3,2] */, 0xf, 0xf,
                                                    false)),
            __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(snd.y), 0xB1, 0xf, 0xf, false)));
        const float4 o = odd ? make_float4(rcv.x, rcv.y, bq.x, bq.y) : make_float4(a.x, a.y, rcv.x, rcv.y);
        const long long row = b2 + zo + 256 * (2 * pp + (odd ? 1 : 0));
        *reinterpret_cast<float4*>(dste + row * kN * kC) = o;
      }
    }
  }
}

// ---------------------------------------------------------------- z pass
// Last inverse pass (z, c2r) with the barrier fused, for blocks of 8
// consecutive z rows (the T layout holds 16 kz of 8 rows contiguously).  Per
// radius the raw FP32 rows are staged in LDS; stage A forms the half-length
// c2r's sequence Z[k] = (a + conj m) + i w^k (a - conj m) (a = X[k], m = X[512
// - k], w = exp(2 pi i / 1024); k = 0 gives the packed slot's (a.x + a.y, a.x -
// a.y)) in FP32 and runs radix 16 (p = 1) on the matrix cores; stage B (p = 16)
// multiplies its inputs by w256^(r k) on the VALU and runs the same radix-16
// matrix; stage C (radix 2, p = 256) leaves row w in wave w, which evaluates
// d2z() (one FP32 FMA per cell), keeps the maximum over the batch and the
// per-radius sums of squares.
//
// The stored intermediates are the y pass's FP32 output, the z spectra T of
// the rows of delta_R: |T| <= max|delta_R| <= 2 max|delta| (the top-hat is a
// positive averaging kernel; the mean is removed), so |Z| <= 8 max|delta| and a
// radix-16 stage grows that by at most 16 sqrt 2.  Stage A's inputs are scaled
// by zscale = a power of two <= 256 / max|delta| (the host's choice from the
// input field), which keeps both stages' FP16 parts below 2^16; the barrier
// undoes the scale exactly.
constexpr int kLz = kN / 2;  // complex points per z row
constexpr int kRowsZ = 8;    // rows per block
constexpr int kZStride = kLz + 2;  // float2 per LDS row: consecutive rows 4 banks apart
// raw rows and stage-B outputs (same buffer): stage-B outputs at pos ^ bit8(pos);
// stage-A outputs: pos ^ bit4(pos) (bank-conflict-free quarter waves, see the
// x pass)
__device__ __forceinline__ int zx_s(int row, int pos) { return row * kZStride + pos; }
__device__ __forceinline__ int zy2_s(int row, int pos) { return row * kZStride + (pos ^ ((pos >> 8) & 1)); }
__device__ __forceinline__ int zy1_s(int row, int pos) { return row * kZStride + (pos ^ ((pos >> 4) & 1)); }

constexpr int kFzThreads = 512;

template <bool RB>
__global__ void __launch_bounds__(kFzThreads) __attribute__((amdgpu_waves_per_eu(4, 4)))
    k_fz_split(const float2* __restrict__ Tbuf, float* __restrict__ zre, float* __restrict__ rbub,
               const RadiusConst* __restrict__ rc, int r0, int nb, double* __restrict__ partial,
               float zscale, const float2* __restrict__ tw_full, const half8_t* __restrict__ mats) {
  constexpr long long nrows = (long long)kN * kN;
  constexpr long long plane = nrows * kLz;
  __shared__ float2 bufX[kRowsZ * kZStride];  // raw rows, then stage B's output
  __shared__ float2 bufY[kRowsZ * kZStride];  // stage A's output
  __shared__ float2 s_w[kLz];                 // exp(2 pi i k / 1024), k < 512 (glue)
  __shared__ float2 s_tw2[kLz / 2];           // exp(2 pi i b / 512) (stage C)
  __shared__ float2 s_twB[256];               // exp(2 pi i m / 256) (stage B)
  __shared__ double red[kRowsZ][kFusedMaxRadii];
  __shared__ BarrierF32 s_bk[kFusedMaxRadii];  // barrier constants of the batch's radii
  __shared__ half8_t s_gD[2][2][64];           // the 16-point DFT, [block h][hi/lo][lane]
  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int g = l >> 4, n = l & 15;
  const int rowA = n & 7, sel = n >> 3;  // MFMA column n = (row, butterfly parity sel)
  if (tid < kRowsZ * kFusedMaxRadii) red[tid / kFusedMaxRadii][tid % kFusedMaxRadii] = 0.0;
  for (int q = tid; q < kLz; q += kFzThreads) s_w[q] = tw_full[q];
  for (int q = tid; q < kLz / 2; q += kFzThreads) s_tw2[q] = tw_full[2 * q];
  for (int q = tid; q < 256; q += kFzThreads) s_twB[q] = tw_full[4 * q];
  if (w < 4) s_gD[w >> 1][w & 1][l] = mats[((w >> 1) * kSplitParts + (w & 1)) * 64 + l];
  // the barrier in FP32, one FMA per cell with the scale folded in exactly:
  // |error| <= (|z| + |A B - 1|) 2^-24, ~2e-6 in z
  if (tid < nb) s_bk[tid] = barrier_f32(rc[r0 + tid].A, rc[r0 + tid].B, 1.0f / zscale);
  const double sc2 = 1.0 / ((double)zscale * (double)zscale);
  // raw-load mapping: a quarter wave loads 16 consecutive kz of one row (128 B):
  // kz = 16 kzb + (t & 15), kzb = ((t >> 4) & 1) + 2 (t >> 8) + 4 e, row (t >> 5) & 7
  const int lrow = (tid >> 5) & 7, lkz = tid & 15, lkzb = ((tid >> 4) & 1) + 2 * (tid >> 8);
  float2 pf[8];
  const long long lane_off = ((long long)lkzb * nrows + lrow) * 16 + lkz;
  auto fetch = [&](long long gbase, int q) {
    const float2* tp = Tbuf + (long long)q * plane + gbase * 16 + lane_off;
#pragma unroll
    for (int e = 0; e < 8; ++e) pf[e] = tp[(long long)(4 * e) * nrows * 16];
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
      int zo = 0;  // opaque zero: the LDS operand reads stay inside the radius loop
      asm volatile("" : "+v"(zo));
#pragma unroll
      for (int e = 0; e < 8; ++e) bufX[zx_s(lrow, 16 * (lkzb + 4 * e) + lkz)] = pf[e];
      __syncthreads();
      // stage A: butterflies b = 4w + 2t + sel of row rowA, inputs Z[b + 32 r]
      // (t loops rolled: each t's operands and accumulators are dead before the next)
#pragma unroll 1
      for (int t = 0; t < 2; ++t) {
        const int b = 4 * w + 2 * t + sel;
        float2 x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = b + 32 * (4 * g + j);
          const float2 a = bufX[zx_s(rowA, k)];
          if (k == 0) {
            x[j] = make_float2((a.x + a.y) * zscale, (a.x - a.y) * zscale);
          } else {
            const float2 mb = cconj(bufX[zx_s(rowA, kLz - k)]);
            const float2 s = cadd(a, mb), dd = csub(a, mb);
            const float2 wd = cmul(s_w[k], dd);
            x[j] = make_float2((s.x - wd.y) * zscale, (s.y + wd.x) * zscale);
          }
        }
        const SplitOp bop = split_b(x);
        const float4_t d0 = mfma_split(SplitOp{s_gD[0][0][l + zo], s_gD[0][1][l + zo]}, bop);
        const float4_t d1 = mfma_split(SplitOp{s_gD[1][0][l + zo], s_gD[1][1][l + zo]}, bop);
        const int p0 = 16 * b + g;
        bufY[zy1_s(rowA, p0)] = make_float2(d0[0], d0[1]);
        bufY[zy1_s(rowA, p0 + 4)] = make_float2(d0[2], d0[3]);
        bufY[zy1_s(rowA, p0 + 8)] = make_float2(d1[0], d1[1]);
        bufY[zy1_s(rowA, p0 + 12)] = make_float2(d1[2], d1[3]);
      }
      __syncthreads();
      if (q + 1 < nb)
        fetch(base, q + 1);
      else if (base + stride < nrows)
        fetch(base + stride, 0);
      // stage B: k = 2w + t, butterflies b' = k + 16 sel of row rowA, inputs
      // b' + 32 r times w256^(r k); outputs to the raw buffer, free since the
      // barrier after stage A
#pragma unroll 1
      for (int t = 0; t < 2; ++t) {
        const int kk = 2 * w + t, bp = kk + 16 * sel;
        float2 x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int r = 4 * g + j;
          x[j] = cmul(bufY[zy1_s(rowA, bp + 32 * r)], s_twB[r * kk]);
        }
        const SplitOp bop = split_b(x);
        const float4_t d0 = mfma_split(SplitOp{s_gD[0][0][l + zo], s_gD[0][1][l + zo]}, bop);
        const float4_t d1 = mfma_split(SplitOp{s_gD[1][0][l + zo], s_gD[1][1][l + zo]}, bop);
        const int p0 = sel * 256 + kk + 16 * g;
        bufX[zy2_s(rowA, p0)] = make_float2(d0[0], d0[1]);
        bufX[zy2_s(rowA, p0 + 64)] = make_float2(d0[2], d0[3]);
        bufX[zy2_s(rowA, p0 + 128)] = make_float2(d1[0], d1[1]);
        bufX[zy2_s(rowA, p0 + 192)] = make_float2(d1[2], d1[3]);
      }
      __syncthreads();
      // stage C: row w, butterflies b'' = l + 64 m; outputs n = b'' and b'' + 256
      float2 v[8];
#pragma unroll
      for (int mm = 0; mm < 4; ++mm) {
        const int b2 = l + 64 * mm;
        const float2 u0 = bufX[zy2_s(w, b2)];
        const float2 u1 = cmul(bufX[zy2_s(w, b2 + 256)], s_tw2[b2]);
        v[mm] = cadd(u0, u1);
        v[mm + 4] = csub(u0, u1);
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
        } else {  // z is never NaN: see barrier_offset
          best[e].x = zmax_nonneg(best[e].x, z0);
          best[e].y = zmax_nonneg(best[e].y, z1);
        }
      }
      const double wsum = (double)wave_sum_f(acc) * sc2;  // 64 lanes in float, then double
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

bool split_supported(int N) { return N == kN; }

size_t split_matrix_bytes(int N) {
  return split_supported(N) ? sizeof(half8_t) * kSplitSets * kSplitParts * 64 : 0;
}

hipError_t launch_split_matrices(void* out, hipStream_t s) {
  hipLaunchKernelGGL(k_split_mats, dim3(kSplitSets), dim3(64), 0, s, static_cast<half8_t*>(out));
  return hipGetLastError();
}

hipError_t launch_fx_split(int N, const float2* S, const float2* D0, const float2* DN, float2* T, float dk,
                           float invN3, const RadiusConst* rc, int r0, int nb, const float2* tw_full,
                           const WTable& wt, const void* mats, hipStream_t s) {
  if (N != kN || nb < 1 || nb > kFxMaxRadii || wt.M < 0 || wt.M > kWTableSplitMax || !mats)
    return hipErrorInvalidValue;
  const long long nblk = (long long)kN * ((kN / 2) / kC) * kNSUB;
  hipLaunchKernelGGL(k_fx_split, dim3((unsigned)nblk), dim3(kFxThreads), sizeof(float4) * wt.M, s, S, D0, DN, T, dk,
                     invN3, rc, r0, nb, tw_full, wt.coef, wt.M, wt.invh, static_cast<const half8_t*>(mats));
  return hipGetLastError();
}

int fz_split_blocks(bool rb) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = rb ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fz_split<true>, kFzThreads, 0)
           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fz_split<false>, kFzThreads, 0);
  long long g = (e == hipSuccess && cus > 0 && per_cu > 0) ? (long long)cus * per_cu : 512;
  if (g > kFzMaxBlocks) g = kFzMaxBlocks;
  const long long groups = (long long)kN * kN / kRowsZ;
  return (int)(g < groups ? g : groups);
}

hipError_t launch_fz_split(int N, const float2* T, float* zre, float* rbub, const RadiusConst* rc, int r0, int nb,
                           double* partial, int nblk, float zscale, const float2* tw_full, const void* mats,
                           hipStream_t s) {
  if (N != kN || nb < 1 || nb > kFusedMaxRadii || nblk < 1 || nblk > kFzMaxBlocks || !mats)
    return hipErrorInvalidValue;
  const half8_t* M = static_cast<const half8_t*>(mats);
  if (rbub)
    hipLaunchKernelGGL(k_fz_split<true>, dim3(nblk), dim3(kFzThreads), 0, s, T, zre, rbub, rc, r0, nb, partial,
                       zscale, tw_full, M);
  else
    hipLaunchKernelGGL(k_fz_split<false>, dim3(nblk), dim3(kFzThreads), 0, s, T, zre, rbub, rc, r0, nb, partial,
                       zscale, tw_full, M);
  return hipGetLastError();
}

}  // namespace prc
