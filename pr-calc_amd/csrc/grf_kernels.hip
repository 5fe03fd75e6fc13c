// GPU Gaussian random field: the synthetic linear density contrast used as
// d2z input (the reference's input producer is the ics program,
// c++/src/grf_ics.C:8-112, which draws rank-dependent drand48 normals; this
// generator keys every normal on (seed, cell) instead).
//
//   white noise  Philox4x32-10 (Salmon et al., SC'11), counter (cell lo, cell hi,
//                0, 0), key (seed lo, seed hi); Box-Muller in double on the first
//                two words, rounded to float
//   colouring    forward r2c, times sqrt(P(k)/V_cell)/N^3 (k = 0 -> 0), c2r
//                (as ConvolveNoise, c++/src/grf_ics.C:34-93)
//   spectrum     P = A k^ns T_BBKS(k)^2, q = k/Gamma, Gamma = Om h exp(-Ob (1 +
//                sqrt(2h)/Om)), k in h/Mpc, A from sigma(8 Mpc/h) = sigma8 by FP64
//                Gauss-Legendre quadrature
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "d2z_device.h"
#include "d2z_kernels.h"
#include "fft_core.h"

namespace prc {

__global__ void k_white_noise(float* __restrict__ out, long long n, unsigned long long seed) {
  const uint2 key = make_uint2((uint32_t)seed, (uint32_t)(seed >> 32));
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n;
       id += (long long)gridDim.x * blockDim.x) {
    const uint4 w = philox4x32_10(make_uint4((uint32_t)id, (uint32_t)((unsigned long long)id >> 32), 0u, 0u), key);
    const double u1 = ((double)(w.x >> 8) + 0.5) * 5.9604644775390625e-08;  // 2^-24
    const double u2 = (double)(w.y >> 8) * 5.9604644775390625e-08;
    out[id] = (float)(sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2));
  }
}

// BBKS transfer function, k in h/Mpc
__device__ double t_bbks(double k, double gamma) {
  const double q = k / gamma;
  if (q == 0.0) return 1.0;
  const double poly = 1 + 3.89 * q + (16.1 * q) * (16.1 * q) + pow(5.46 * q, 3.0) + pow(6.71 * q, 4.0);
  return log(1 + 2.34 * q) / (2.34 * q) * pow(poly, -0.25);
}

__device__ double w_tophat_d(double x) {
  if (x < 1e-3) return 1 - x * x / 10;
  return 3 * (sin(x) - x * cos(x)) / (x * x * x);
}

// sigma^2(8 Mpc/h) for A = 1: integral over ln k in [ln 1e-6, ln 1e4] of
// k^(3+ns) T^2 W(8k)^2 / (2 pi^2); one block, 8-point Gauss-Legendre on 4096 panels.
__global__ void __launch_bounds__(256) k_grf_norm(GrfSpectrum sp, double* __restrict__ out) {
  const double gx[8] = {-0.9602898564975362, -0.7966664774136267, -0.525532409916329, -0.18343464249564978,
                        0.18343464249564978, 0.525532409916329,   0.7966664774136267, 0.9602898564975362};
  const double gw[8] = {0.10122853629037669, 0.22238103445337434, 0.31370664587788705, 0.36268378337836177,
                        0.36268378337836177, 0.31370664587788705, 0.22238103445337434, 0.10122853629037669};
  const double a = log(1e-6), b = log(1e4);
  constexpr int NP = 4096;
  const double hw = (b - a) / NP * 0.5;
  double acc = 0.0;
  for (int p = threadIdx.x; p < NP; p += blockDim.x) {
    const double mid = a + hw * (2 * p + 1);
    double s = 0.0;
    for (int q = 0; q < 8; ++q) {
      const double k = exp(mid + hw * gx[q]);
      const double t = t_bbks(k, sp.gamma);
      const double w = w_tophat_d(8.0 * k);
      s += gw[q] * pow(k, 3 + sp.ns) * t * t * w * w;
    }
    acc += s * hw;
  }
  acc /= 2 * M_PI * M_PI;
  __shared__ double sh[256];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = (double)sp.sigma8 * sp.sigma8 / sh[0];
}

// amplitude sqrt(P(k)/V_cell)/N^3 of grid mode with |k|^2 = m dk^2
__device__ __forceinline__ float grf_amp(int m, const GrfSpectrum& sp, double A) {
  if (m == 0) return 0.0f;
  const double k = sqrt((double)m) * sp.dk;
  const double t = t_bbks(k, sp.gamma);
  const double P = A * pow(k, sp.ns) * t * t;
  return (float)(sqrt(P / sp.vcell) * sp.invN3);
}

// colour the half spectrum S in place; the kz = 0 slot is rebuilt from the
// side planes as a0 D0 + i aN DN (the packed pair of the z transform)
template <int C>
__global__ void k_colour(float2* __restrict__ S, const float2* __restrict__ D0, const float2* __restrict__ DN,
                         int N, GrfSpectrum sp, const double* __restrict__ Aptr) {
  const double A = Aptr[0];
  const long long Nh = N / 2;
  const long long total = (long long)N * N * Nh;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    const SIndex k = s_index<C>(id, N);
    const int mxy = fsq(k.kx, N) + fsq(k.ky, N);
    if (k.kz != 0) {
      const float a = grf_amp(mxy + k.kz * k.kz, sp, A);
      const float2 v = S[id];
      S[id] = make_float2(v.x * a, v.y * a);
    } else {
      const long long pl = (long long)k.kx * N + k.ky;
      const float a0 = grf_amp(mxy, sp, A), an = grf_amp(mxy + (int)(Nh * Nh), sp, A);
      const float2 p = D0[pl], q = DN[pl];
      S[id] = make_float2(a0 * p.x - an * q.y, a0 * p.y + an * q.x);
    }
  }
}

hipError_t launch_white_noise(float* out, long long n, unsigned long long seed, hipStream_t s) {
  long long g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  hipLaunchKernelGGL(k_white_noise, dim3((unsigned)g), dim3(256), 0, s, out, n, seed);
  return hipGetLastError();
}

hipError_t launch_grf_norm(const GrfSpectrum& sp, double* A, hipStream_t s) {
  hipLaunchKernelGGL(k_grf_norm, dim3(1), dim3(256), 0, s, sp, A);
  return hipGetLastError();
}

hipError_t launch_colour(int N, float2* S, const float2* D0, const float2* DN, const GrfSpectrum& sp,
                         const double* A, hipStream_t s) {
  const long long n = (long long)N * N * (N / 2);
  long long g = (n + 255) / 256;
  if (g > (1 << 20)) g = 1 << 20;
  const int C = block_width(N);
  if (C == 16)
    hipLaunchKernelGGL((k_colour<16>), dim3((unsigned)g), dim3(256), 0, s, S, D0, DN, N, sp, A);
  else if (C == 8)
    hipLaunchKernelGGL((k_colour<8>), dim3((unsigned)g), dim3(256), 0, s, S, D0, DN, N, sp, A);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace prc
