// History correction of the zreion field: bubble mean free paths from random
// skewers (GetBubbleMFP, c++/src/bubblehistory_d2z.C:82-242) and the remap of
// every cell through the corrected history (zreioncorrect and CorrectHistory,
// c++/src/bubblehistory_d2z.C:270-343).
//
//   skewers  one thread per (skewer, ionized fraction): start cells drawn from
//            Philox4x32-10 (counter (skewer, fraction, draw, 0), key (seed lo,
//            seed hi)) until zreion > zcur with j, k < N - 2, a bounded number
//            of draws; walks along +y and +z (periodic) to the first cell with
//            zreion < zcur or back to the start; distance log-uniform in
//            [m - 1, m] cells ([1/2, 1] for m = 1) from counter (skewer,
//            fraction, 0xffffffff, 0); distances and their logs in FP64
//   zvz      one thread per history bin, the corrected-history table in LDS,
//            first bracketing interval as the reference's linear search
//   remap    zreion <- zvz[bin(zreion)], bin = (int)((z - 49.9)/dzhistory)
//            clamped to the table (the reference indexes out of bounds for
//            z < 0.01 and z > 49.9)
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "d2z_device.h"
#include "d2z_kernels.h"

namespace prc {
namespace {

constexpr int kSkewerThreads = 256;

__device__ __forceinline__ float u24(uint32_t w) { return (float)(w >> 8) * 0x1p-24f; }

// (int)(u N) with the reference's wrap (bubblehistory_d2z.C:140-150)
__device__ __forceinline__ int cell_of(float u, int n) {
  int c = (int)(u * (float)n);
  if (c >= n) c -= n;
  if (c < 0) c += n;
  return c;
}

// exp(clow + (chigh - clow) frac) cells, clow = log(m - 1) (log 1/2 for m = 1),
// chigh = log m (bubblehistory_d2z.C:185-189,223-227), in FP64 without contraction
// so that the host restatement rounds identically
__device__ double skewer_distance(int m, float frac, double cell) {
#pragma clang fp contract(off)
  const double clow = m > 1 ? log((double)m - 1.0) : log(0.5);
  const double chigh = log((double)m);
  return exp(clow + (chigh - clow) * (double)frac) * cell;
}

__global__ void __launch_bounds__(kSkewerThreads)
    k_skewers(const float* __restrict__ zre, const float* __restrict__ zcur, SkewerParams p,
              double* __restrict__ sums, unsigned int* __restrict__ counts, int* __restrict__ fail) {
#pragma clang fp contract(off)
  __shared__ unsigned int h[kNMfp];
  __shared__ double ws[kSkewerThreads / 64];
  const int t = threadIdx.x, ix = blockIdx.y;
  for (int b = t; b < kNMfp; b += kSkewerThreads) h[b] = 0u;
  __syncthreads();
  const int s = blockIdx.x * kSkewerThreads + t;
  const int N = p.N;
  const long long NN = (long long)N * N;
  double dsum = 0.0;
  if (s < p.nskewers) {
    const float zc = zcur[ix];
    const uint2 key = make_uint2(p.seed_lo, p.seed_hi);
    int i0 = 0, j0 = 0, k0 = 0;
    bool found = false;
    for (uint32_t a = 0; a < (uint32_t)kMaxStartDraws && !found; ++a) {
      // another skewer already ran out of draws: the call fails, stop early
      if ((a & 255u) == 255u && __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      const uint4 w = philox4x32_10(make_uint4((uint32_t)s, (uint32_t)ix, a, 0u), key);
      i0 = cell_of(u24(w.x), N);
      j0 = cell_of(u24(w.y), N);
      k0 = cell_of(u24(w.z), N);
      found = zre[i0 * NN + (long long)j0 * N + k0] > zc && j0 < N - 2 && k0 < N - 2;
    }
    if (!found) {
      atomicOr(fail, 1);
    } else {
      const uint4 wf = philox4x32_10(make_uint4((uint32_t)s, (uint32_t)ix, 0xffffffffu, 0u), key);
      const float* line = zre + i0 * NN;
      // +y (bubblehistory_d2z.C:159-183), then +z (:196-221): at most N steps each
      int m = 0, j = j0;
      for (;;) {
        j = (j + 1 == N) ? 0 : j + 1;
        ++m;
        if (line[(long long)j * N + k0] < zc || j == j0) break;
      }
      const double dy = skewer_distance(m, u24(wf.x), p.cell);
      m = 0;
      int k = k0;
      const float* row = line + (long long)j0 * N;
      for (;;) {
        k = (k + 1 == N) ? 0 : k + 1;
        ++m;
        if (row[k] < zc || k == k0) break;
      }
      const double dz = skewer_distance(m, u24(wf.y), p.cell);
      dsum = dy + dz;
      // bins of log distance (bubblehistory_d2z.C:191-194,229-232)
      const double ld[2] = {log(dy), log(dz)};
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int b = (int)((ld[q] - (double)p.lmin) / (double)p.dmfp);
        if (b >= 0 && b < kNMfp && ld[q] < (double)p.lmax && ld[q] > (double)p.lmin) atomicAdd(&h[b], 1u);
      }
    }
  }
  const double wsum = wave_sum(dsum);
  if ((t & 63) == 0) ws[t >> 6] = wsum;
  __syncthreads();
  if (t == 0) {
    double tot = 0.0;
    for (int w = 0; w < kSkewerThreads / 64; ++w) tot += ws[w];
    if (tot != 0.0) atomicAdd(&sums[ix], tot);
  }
  for (int b = t; b < kNMfp; b += kSkewerThreads)
    if (h[b]) atomicAdd(&counts[ix * kNMfp + b], h[b]);
}

// zreioncorrect (bubblehistory_d2z.C:270-299) for every history bin; -1 where
// the bin is 0 or its ionized fraction is 0 (CorrectHistory, :305-314); NaN
// where no interval brackets x (the reference exits there)
__global__ void __launch_bounds__(256)
    k_zvz(const float* __restrict__ history, const float* __restrict__ chist, float dzch, float* __restrict__ zvz) {
#pragma clang fp contract(off)
  __shared__ float c[kNCHistory];
  for (int q = threadIdx.x; q < kNCHistory; q += 256) c[q] = chist[q];
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kNHistory) return;
  const float x = history[i];
  float out = -1.0f;
  if (i > 0 && x != 0.0f) {
    if (x < c[0]) {
      out = 49.9f;
    } else if (x > c[kNCHistory - 1]) {
      out = 0.01f;
    } else {
      out = __builtin_nanf("");
      for (int q = 0; q < kNCHistory - 1; ++q) {
        if (x >= c[q] && x <= c[q + 1]) {
          const float f = (x - c[q]) / (c[q + 1] - c[q]);
          const double z0 = 49.9 + (double)((float)q * dzch);
          const double z1 = 49.9 + (double)((float)(q + 1) * dzch);
          out = (float)((double)(1.0f - f) * z0 + (double)f * z1);
          break;
        }
      }
    }
  }
  zvz[i] = out;
}

__device__ __forceinline__ float remap_one(float z, const float* __restrict__ zvz, double dz) {
  const double q = ((double)z - 49.9) / dz;
  const int b = q >= (double)kNHistory ? kNHistory - 1 : (q > 0.0 ? (int)q : 0);
  return zvz[b];
}

__global__ void __launch_bounds__(256)
    k_remap(float* __restrict__ zre, long long n, const float* __restrict__ zvz, double dz) {
  const long long n4 = n >> 2;
  float4* z4 = reinterpret_cast<float4*>(zre);
  const long long stride = (long long)gridDim.x * 256;
  for (long long id = (long long)blockIdx.x * 256 + threadIdx.x; id < n4; id += stride) {
    float4 v = z4[id];
    v.x = remap_one(v.x, zvz, dz);
    v.y = remap_one(v.y, zvz, dz);
    v.z = remap_one(v.z, zvz, dz);
    v.w = remap_one(v.w, zvz, dz);
    z4[id] = v;
  }
  for (long long id = 4 * n4 + (long long)blockIdx.x * 256 + threadIdx.x; id < n; id += stride)
    zre[id] = remap_one(zre[id], zvz, dz);
}

}  // namespace

hipError_t launch_skewers(const float* zre, const float* zcur, const SkewerParams& p, double* sums,
                          unsigned int* counts, int* fail, hipStream_t s) {
  if (p.N < 3 || p.nskewers < 1) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((p.nskewers + kSkewerThreads - 1) / kSkewerThreads), kNXMfp);
  hipLaunchKernelGGL(k_skewers, grid, dim3(kSkewerThreads), 0, s, zre, zcur, p, sums, counts, fail);
  return hipGetLastError();
}

hipError_t launch_zvz(const float* history, const float* chist, float dzch, float* zvz, hipStream_t s) {
  hipLaunchKernelGGL(k_zvz, dim3((kNHistory + 255) / 256), dim3(256), 0, s, history, chist, dzch, zvz);
  return hipGetLastError();
}

hipError_t launch_remap(float* zre, long long n, const float* zvz, double dz, hipStream_t s) {
  long long g = (n / 4 + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_remap, dim3((unsigned)g), dim3(256), 0, s, zre, n, zvz, dz);
  return hipGetLastError();
}

}  // namespace prc
