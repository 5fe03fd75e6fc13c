// Set-up and element-wise kernels of the delta2zreion path (gfx950): twiddle
// tables, the FP64 scalar set-up (sigma(R) quadrature, erfc^-1 root, growth
// factors), the plain path's barrier and moments, the history histogram.
//
// Reference: c++/src/delta2zreion_d2z.C:12-38,97-146,206-216,
// c++/src/cosmology.C:174-309, c++/src/math.C:23-61,
// c++/src/arrayoperations.C:4-35, c++/src/bubblehistory_d2z.C:5-41.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "d2z_device.h"
#include "d2z_kernels.h"
#include "fft_core.h"

namespace prc {

// --------------------------------------------------------------- twiddles
// tw[m] = exp(+2 pi i m / L), m < L, from FP64 sincospi rounded to float.
__global__ void k_twiddles(float2* __restrict__ tw, int L) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < L) {
    double s, c;
    sincospi(2.0 * (double)m / (double)L, &s, &c);
    tw[m] = make_float2((float)c, (float)s);
  }
}

// plain path: strict running maximum over a real field (delta2zreion_d2z.C:106-146)
__global__ void k_barrier(const float* __restrict__ dR, float* __restrict__ zre, float* __restrict__ rbub,
                          long long total, const RadiusConst* __restrict__ rc, int r) {
  const double A = rc[r].A;
  const double B = rc[r].B;
  const float R = rc[r].R;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    const float z = barrier_z(dR[id], A, B);
    if (z > zre[id]) {
      zre[id] = z;
      if (rbub) rbub[id] = R;
    }
  }
}

// ------------------------------------------ sigma_grid partial sums (plain)
// Per-block FP64 sums of delta and delta^2 (GetMeanSigma, arrayoperations.C:4-35).

__global__ void __launch_bounds__(256) k_moments(const float* __restrict__ x, long long total,
                                                 double2* __restrict__ partial) {
  double s1 = 0.0, s2 = 0.0;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    const double d = (double)x[id];
    s1 += d;
    s2 += d * d;
  }
  __shared__ double sh[2][4];
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) {
    sh[0][w] = s1;
    sh[1][w] = s2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int k = 0; k < (int)(blockDim.x / 64); ++k) {
      a += sh[0][k];
      b += sh[1][k];
    }
    partial[blockIdx.x] = make_double2(a, b);
  }
}

// Fixed-order reduction of per-block partials -> sigma_grid = sqrt(sum d^2 / N^3)
__global__ void k_moments_finish(const double2* __restrict__ partial, int nblk, int nrad, double invN3,
                                 float* __restrict__ sigma_grid_out, int r0) {
  const int r = blockIdx.x;
  if (r >= nrad) return;
  double a = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 64) a += partial[(long long)r * nblk + k].y;
  a = wave_sum(a);
  if (threadIdx.x == 0) sigma_grid_out[r0 + r] = (float)sqrt(a * invN3);
}

// ------------------------------------------------ input check
// flag = 1 if any cell is NaN or Inf (plain vector stores of the same value).
__global__ void k_nonfinite(const float* __restrict__ x, long long n, int* __restrict__ flag) {
  int bad = 0;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n;
       id += (long long)gridDim.x * blockDim.x)
    bad |= !isfinite(x[id]);
  if (bad) *flag = 1;
}

// max |x| as the bit pattern of a non-negative float (ordered as unsigned)
__global__ void k_maxabs(const float* __restrict__ x, long long n, unsigned int* __restrict__ out) {
  float m = 0.0f;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n;
       id += (long long)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(x[id]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

// ------------------------------------------------ history histogram (GPU)
// bin = (int)((z - 49.9)/dzhistory) in double (bubblehistory_d2z.C:17-28);
// integer counts.  A first pass finds the lowest and highest occupied bin; then
// block (chunk, w) counts the bins of window [lo + w W, lo + (w + 1) W) in an LDS
// histogram (LDS atomics) and adds its non-zero bins to the global counts.
// Windows past the highest bin return at once, so a field whose bins span
// fewer than W (the usual case: zreion between ~5 and ~20 spans ~30000) is read
// twice in all.
constexpr int kHistWindow = 36000;  // bins per window: 144 KB of LDS
constexpr int kHistWindows = (kNHistory + kHistWindow - 1) / kHistWindow;

__device__ __forceinline__ int history_bin(float z, double dz) {
  const double q = ((double)z - 49.9) / dz;
  return (q > -1.0 && q < (double)kNHistory) ? (int)q : -1;  // NaN: -1
}

__global__ void __launch_bounds__(256) k_history_range(const float* __restrict__ zre, long long total,
                                                       int* __restrict__ range, double dz) {
  int lo = kNHistory, hi = -1;
  const long long n4 = total >> 2;
  const float4* z4 = reinterpret_cast<const float4*>(zre);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n4; id += stride) {
    const float4 v = z4[id];
    const int b[4] = {history_bin(v.x, dz), history_bin(v.y, dz), history_bin(v.z, dz), history_bin(v.w, dz)};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (b[q] >= 0) {
        lo = min(lo, b[q]);
        hi = max(hi, b[q]);
      }
  }
  for (long long id = 4 * n4 + (long long)blockIdx.x * blockDim.x + threadIdx.x; id < total; id += stride) {
    const int b = history_bin(zre[id], dz);
    if (b >= 0) {
      lo = min(lo, b);
      hi = max(hi, b);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, __shfl_xor(lo, off, 64));
    hi = max(hi, __shfl_xor(hi, off, 64));
  }
  if ((threadIdx.x & 63) == 0 && hi >= 0) {
    atomicMin(&range[0], lo);
    atomicMax(&range[1], hi);
  }
}

__global__ void __launch_bounds__(1024) k_history(const float* __restrict__ zre, long long total,
                                                  unsigned int* __restrict__ counts, double dz,
                                                  const int* __restrict__ range) {
  __shared__ unsigned int h[kHistWindow];
  if (range[1] < 0) return;  // no cell in any bin
  const int lo = range[0] + (int)blockIdx.y * kHistWindow;
  if (lo > range[1]) return;  // window past the highest bin (block-uniform)
  for (int b = threadIdx.x; b < kHistWindow; b += blockDim.x) h[b] = 0u;
  __syncthreads();
  const long long n4 = total >> 2;
  const float4* z4 = reinterpret_cast<const float4*>(zre);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n4; id += 2 * stride) {
    const float4 v = z4[id];
    const float4 u = id + stride < n4 ? z4[id + stride] : make_float4(NAN, NAN, NAN, NAN);
    const int b[8] = {history_bin(v.x, dz), history_bin(v.y, dz), history_bin(v.z, dz), history_bin(v.w, dz),
                      history_bin(u.x, dz), history_bin(u.y, dz), history_bin(u.z, dz), history_bin(u.w, dz)};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int r = b[q] - lo;
      if (b[q] >= 0 && r >= 0 && r < kHistWindow) atomicAdd(&h[r], 1u);
    }
  }
  for (long long id = 4 * n4 + (long long)blockIdx.x * blockDim.x + threadIdx.x; id < total; id += stride) {
    const int b = history_bin(zre[id], dz), r = b - lo;
    if (b >= 0 && r >= 0 && r < kHistWindow) atomicAdd(&h[r], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kHistWindow; b += blockDim.x)
    if (h[b]) atomicAdd(&counts[lo + b], h[b]);
}

// =================================================== FP64 scalar set-up
// sigma^2(R) = pnorm^2 * integral over ln k of k^(3+ns) T_EH(k)^2 W(kR)^2
// (cosmology.C:214-309) by composite 8-point Gauss-Legendre on 4096 panels
// over [ln(1e-5/R), ln(1e5/R)]: relative error ~1e-12, below the 1e-7 the
// reference requests from QAGS.
struct EHd {
  double om, ob, hh, an;
};

__device__ double transeh_d(double k, const EHd& c) {
  const double om = c.om, ob = c.ob, hh = c.hh;
  const double alpha = 1 - 0.328 * log(431 * om * hh * hh) * (ob / om) +
                       0.38 * log(22.3 * om * hh * hh) * (ob / om) * (ob / om);
  const double sound = 44.5 * log(9.83 / (om * hh * hh)) / sqrt(1 + 10 * pow(ob * hh * hh, 3. / 4.));
  const double ks = 0.43 * k * sound;
  const double shape = om * hh * (alpha + (1 - alpha) / (1 + ks * ks * ks * ks));
  const double q = (k / hh) * ((2.728 / 2.7) * (2.728 / 2.7)) / shape;
  const double lg = log(2 * 2.718281828459045235 + 1.8 * q);
  return lg / (lg + (14.2 + 731 / (1 + 62.5 * q)) * q * q);
}

__device__ double sgm2k_d(double lk, double R, const EHd& c) {
  const double k = exp(lk);
  const double kR = k * R;
  const double Pin = pow(k, 3 + c.an);
  const double T = transeh_d(k, c);
  const double W = 3 * (sin(kR) - kR * cos(kR)) / (kR * kR * kR);
  return Pin * T * T * W * W;
}

__constant__ double kGLx[8] = {-0.9602898564975362, -0.7966664774136267, -0.525532409916329,
                               -0.18343464249564978, 0.18343464249564978, 0.525532409916329,
                               0.7966664774136267, 0.9602898564975362};
__constant__ double kGLw[8] = {0.10122853629037669, 0.22238103445337434, 0.31370664587788705,
                               0.36268378337836177, 0.36268378337836177, 0.31370664587788705,
                               0.22238103445337434, 0.10122853629037669};

// float radius schedule R_r = 10^(log10 R0 + r dScale) (delta2zreion_d2z.C:36-38,55)
__device__ float radius_of(const SetupParams& p, int r) {
#pragma clang fp contract(off)
  const float box = p.BoxSize / p.h;
  const float Rinitial = (box * p.Rmin) / (float)p.N;
  const float Rfinal = p.Rmax / p.h;
  // log10 of a float, correctly rounded (double evaluation then float)
  const float l0 = (float)log10((double)Rinitial), l1 = (float)log10((double)Rfinal);
  const float dScale = (l1 - l0) / (float)(p.Nscales - 1);
  const float e = l0 + (float)r * dScale;
  return (float)pow(10.0, (double)e);
}

// R of cdmehint from a float mass: pow(3*mass/fourpi/dens, 1/3) (cosmology.C:221-231)
__device__ double cdm_radius(const SetupParams& p, float mass) {
#pragma clang fp contract(off)
  const float rhoc = 2.775e11f;
  const float pi = 3.141592654f;
  const float fourpi = 4 * pi;
  const double dens = (double)(((p.Omegam * rhoc) * p.h) * p.h);
  const float num = (3.0f * mass) / fourpi;
  return pow((double)num / dens, 1. / 3.);
}

// One block per integral: j = 0 sigma8 normalisation (R = 8/h), j = 1 the
// minimum mass, j >= 2 radius j-2.
__global__ void __launch_bounds__(256) k_sigma_integrals(SetupParams p, double* __restrict__ integ,
                                                         float* __restrict__ Rout, double* __restrict__ Rint) {
  const int j = blockIdx.x;
  double R;
  if (j == 0) {
    R = (double)(8.0f / p.h);
  } else if (j == 1) {
    R = cdm_radius(p, p.M_min);
  } else {
    const float Rr = radius_of(p, j - 2);
    const float rho0 = (float)(2.775e11 * (double)p.Omegam * (double)p.h * (double)p.h);
    const double Rd = (double)Rr;
    const float mass = (float)(4 * M_PI / 3 * (double)rho0 * Rd * Rd * Rd);
    R = cdm_radius(p, mass);
    if (threadIdx.x == 0) Rout[j - 2] = Rr;
  }
  if (threadIdx.x == 0) Rint[j] = R;
  const EHd c{(double)p.Omegam, (double)p.Omegab, (double)p.h, (double)p.ns};
  const double a = log(1e-5 / R), b = log(1e5 / R);
  constexpr int NP = 4096;
  const double hw = (b - a) / NP * 0.5;
  double acc = 0.0;
  for (int pnl = threadIdx.x; pnl < NP; pnl += blockDim.x) {
    const double mid = a + hw * (2 * pnl + 1);
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) s += kGLw[q] * sgm2k_d(mid + hw * kGLx[q], R, c);
    acc += s * hw;
  }
  __shared__ double sh[4];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    integ[j] = s;
  }
}

// growth factor D(z), float Carroll-Press-Turner fit (cosmology.C:174-212)
__device__ float growth_f(float z, float omegam, float omegal, float w) {
#pragma clang fp contract(off)
  const float x = 1 + z;
  const float x2 = x * x;
  const float x3 = x * x * x;
  const float x3w = powf(x3, w);
  const float curv = (1 - omegam) - omegal;
  const float omega = omegam * x3 / ((omegam * x3 + curv * x2) + omegal);
  const float lambda = omegal * x3 * x3w / ((omegam * x3 + curv * x2) + omegal * x3 * x3w);
  float g = (float)(2.5 * (double)omega /
                    (pow((double)omega, 4. / 7.) - (double)lambda +
                     (double)((1 + omega / 2) * (1 + lambda / 70))));
  float g0 = (float)(2.5 * (double)omegam /
                     (pow((double)omegam, 4. / 7.) - (double)omegal +
                      (double)((1 + omegam / 2) * (1 + omegal / 70))));
  const double wd = (double)w;
  const double a1 = 0.255 + 0.305 * wd + 0.0027 / wd;
  const double a2 = 0.366 + 0.266 * wd - 0.0700 / wd;
  const float t = (float)(-a1 * (double)(1 - omega) - a2 * (double)logf(omega));
  const float t0 = (float)(-a1 * (double)(1 - omegam) - a2 * (double)logf(omegam));
  g *= powf(-w, t);
  g0 *= powf(-w, t0);
  return (g / x) / g0;
}

// Brent's method on erfc(t) - a over [0,100] until the bracket passes
// |hi - lo| < 1e-4 * min(|lo|,|hi|); returns the lower end (math.C:23-61).
__device__ double inverse_erfc_lo(double a) {
  const double eps = 2.220446049250313e-16;
  double x_lo = 0.0, x_hi = 100.0;
  double A = x_lo, B = x_hi;
  double fa = erfc(A) - a, fb = erfc(B) - a;
  double Cc = B, fc = fb;
  double d = B - A, e = B - A;
  double lo = x_lo, hi = x_hi;
  for (int it = 0; it < 100; ++it) {
    bool ac_equal = false;
    if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) {
      ac_equal = true;
      Cc = A;
      fc = fa;
      d = e = B - A;
    }
    if (fabs(fc) < fabs(fb)) {
      ac_equal = true;
      A = B;
      B = Cc;
      Cc = A;
      fa = fb;
      fb = fc;
      fc = fa;
    }
    const double tol = 0.5 * eps * fabs(B);
    const double m = 0.5 * (Cc - B);
    if (fb == 0) return B;
    if (fabs(m) <= tol) return (B < Cc) ? B : Cc;
    if (fabs(e) < tol || fabs(fa) <= fabs(fb)) {
      d = e = m;
    } else {
      double pp, qq;
      const double s = fb / fa;
      if (ac_equal) {
        pp = 2 * m * s;
        qq = 1 - s;
      } else {
        double q = fa / fc, r = fb / fc;
        pp = s * (2 * m * q * (q - r) - (B - A) * (r - 1));
        qq = (q - 1) * (r - 1) * (s - 1);
      }
      if (pp > 0)
        qq = -qq;
      else
        pp = -pp;
      if (2 * pp < fmin(3 * m * qq - fabs(tol * qq), fabs(e * qq))) {
        e = d;
        d = pp / qq;
      } else {
        d = e = m;
      }
    }
    A = B;
    fa = fb;
    B += (fabs(d) > tol) ? d : (m > 0 ? tol : -tol);
    fb = erfc(B) - a;
    if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) Cc = A;
    lo = (B < Cc) ? B : Cc;
    hi = (B < Cc) ? Cc : B;
    double min_abs = 0.0;
    if ((lo > 0 && hi > 0) || (lo < 0 && hi < 0)) min_abs = fmin(fabs(lo), fabs(hi));
    if (fabs(hi - lo) < 1e-4 * min_abs) return lo;
  }
  return lo;
}

// Scalars: pnorm, erfizeta, sigma_min, D(zInit), D(20), A and the per-radius
// barrier constants (delta2zreion_d2z.C:17-26,97-100,206-216; main_d2z.C:57-64).
__global__ void k_setup_finish(SetupParams p, const double* __restrict__ integ, const float* __restrict__ Rr,
                               SetupScalars* __restrict__ sc, RadiusConst* __restrict__ rc,
                               float* __restrict__ sigma_out) {
#pragma clang fp contract(off)
  __shared__ float s_pnorm, s_smin, s_erfz, s_A;
  if (threadIdx.x == 0) {
    const float pnorm = (float)((double)p.sigma8 / sqrt(integ[0]));
    const double pn = (double)pnorm;
    const float sgm2min = (float)(integ[1] * pn * pn);
    const float smin = sqrtf(sgm2min);
    const float x = (float)(1.0 / (double)p.zeta);
    const float erfz = (float)inverse_erfc_lo((double)x);
    const float DInit = growth_f(p.zInit, p.Omegam, p.Omegal, p.w);
    const float zHigh = 20.0f;
    const float DHigh = growth_f(zHigh, p.Omegam, p.Omegal, p.w);
    const float A = (1 + zHigh) * DHigh / 1.686f;
    sc->pnorm = pnorm;
    sc->sigma_min = smin;
    sc->erfizeta = erfz;
    sc->DInit = DInit;
    sc->DHigh = DHigh;
    sc->A = A;
    s_pnorm = pnorm;
    s_smin = smin;
    s_erfz = erfz;
    s_A = A;
  }
  __syncthreads();
  const double pn = (double)s_pnorm;
  for (int r = threadIdx.x; r < p.Nscales; r += blockDim.x) {
    const float sgm2 = (float)(integ[2 + r] * pn * pn);
    const float sigma = sqrtf(sgm2);
    const float dd = s_smin * s_smin - sigma * sigma;
    RadiusConst c;
    c.R = Rr[r];
    c.sigma = sigma;
    c.A = (double)s_A;
    c.B = barrier_offset(dd, s_erfz);
    rc[r] = c;
    sigma_out[r] = sigma;
  }
}

// zero fill (graph-safe replacement for a memset)
__global__ void k_fill(float* __restrict__ p, long long n, float v) {
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n;
       id += (long long)gridDim.x * blockDim.x)
    p[id] = v;
}

// sum of the per-block partials of one radius -> sigma_grid
__global__ void k_sigma_finish(const double* __restrict__ partial, int nblk, int nrad, double invN3,
                               float* __restrict__ sigma_grid_out) {
  const int r = blockIdx.x;
  if (r >= nrad) return;
  double a = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 64) a += partial[(long long)r * nblk + k];
  a = wave_sum(a);
  if (threadIdx.x == 0) sigma_grid_out[r] = (float)sqrt(a * invN3);
}

// ============================================================ launchers
static inline unsigned grid1(long long n, int bs, long long cap = 1 << 20) {
  long long g = (n + bs - 1) / bs;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (unsigned)g;
}

hipError_t launch_twiddles(float2* tw, int L, hipStream_t s) {
  hipLaunchKernelGGL(k_twiddles, dim3((L + 255) / 256), dim3(256), 0, s, tw, L);
  return hipGetLastError();
}

hipError_t launch_barrier(const float* dR, float* zre, float* rbub, long long total, const RadiusConst* rc, int r,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_barrier, dim3(grid1(total, 256)), dim3(256), 0, s, dR, zre, rbub, total, rc, r);
  return hipGetLastError();
}

hipError_t launch_moments(const float* x, long long total, double2* partial, int nblk, hipStream_t s) {
  hipLaunchKernelGGL(k_moments, dim3(nblk), dim3(256), 0, s, x, total, partial);
  return hipGetLastError();
}

hipError_t launch_moments_finish(const double2* partial, int nblk, int nrad, double invN3, float* sigma_grid_out,
                                 int r0, hipStream_t s) {
  hipLaunchKernelGGL(k_moments_finish, dim3(nrad), dim3(64), 0, s, partial, nblk, nrad, invN3, sigma_grid_out,
                     r0);
  return hipGetLastError();
}

hipError_t launch_history(const float* zre, long long total, unsigned int* counts, double dz, int* range,
                          hipStream_t s) {
  // range = {0x7f7f7f7f (above every bin), -1} before the min/max pass
  hipError_t e = hipMemsetAsync(range, 0x7f, sizeof(int), s);
  if (e == hipSuccess) e = hipMemsetAsync(range + 1, 0xff, sizeof(int), s);
  if (e != hipSuccess) return e;
  long long g = (total / 4 + 255) / 256;
  if (g > 2048) g = 2048;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_history_range, dim3((unsigned)g), dim3(256), 0, s, zre, total, range, dz);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  long long gh = (total / 8 + 1023) / 1024;
  if (gh > 256) gh = 256;
  if (gh < 1) gh = 1;
  hipLaunchKernelGGL(k_history, dim3((unsigned)gh, kHistWindows), dim3(1024), 0, s, zre, total, counts, dz, range);
  return hipGetLastError();
}

hipError_t launch_setup(const SetupParams& p, double* integ, float* Rr, double* Rint, SetupScalars* sc,
                        RadiusConst* rc, float* sigma_out, hipStream_t s) {
  hipLaunchKernelGGL(k_sigma_integrals, dim3(p.Nscales + 2), dim3(256), 0, s, p, integ, Rr, Rint);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_setup_finish, dim3(1), dim3(64), 0, s, p, integ, Rr, sc, rc, sigma_out);
  return hipGetLastError();
}

hipError_t launch_fill(float* p, long long n, float v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(grid1(n, 256, 16384)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_sigma_finish(const double* partial, int nblk, int nrad, double invN3, float* sigma_grid_out,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_sigma_finish, dim3(nrad), dim3(64), 0, s, partial, nblk, nrad, invN3, sigma_grid_out);
  return hipGetLastError();
}

// ------------------------------------------------- W(x) interpolation table
// W and W' in FP64: Taylor series below x = 1/2 (terms to x^12, truncation
// below 1e-16), closed forms above.
__device__ double w_exact(double x) {
  if (x < 0.5) {
    const double x2 = x * x;
    return 1.0 - x2 * (1.0 / 10 - x2 * (1.0 / 280 - x2 * (1.0 / 15120 - x2 * (1.0 / 1330560 - x2 * (1.0 / 172972800 -
                                                                                                    x2 / 31135104000.0)))));
  }
  double sn, cs;
  sincos(x, &sn, &cs);
  return 3.0 * (sn - x * cs) / (x * x * x);
}
__device__ double w_deriv(double x) {
  if (x < 0.5) {
    const double x2 = x * x;
    return x * (-1.0 / 5 + x2 * (1.0 / 70 - x2 * (1.0 / 2520 - x2 * (1.0 / 166320 - x2 / 17297280.0))));
  }
  double sn, cs;
  sincos(x, &sn, &cs);
  return 3.0 * (x * x * sn - 3.0 * sn + 3.0 * x * cs) / (x * x * x * x);
}
__global__ void k_wtable(float4* __restrict__ coef, int M, double h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const double x0 = i * h, x1 = (i + 1) * h;
  const double w0 = w_exact(x0), w1 = w_exact(x1);
  const double d0 = w_deriv(x0) * h, d1 = w_deriv(x1) * h;
  coef[i] = make_float4((float)w0, (float)d0, (float)(3.0 * (w1 - w0) - 2.0 * d0 - d1),
                        (float)(2.0 * (w0 - w1) + d0 + d1));
}

int wtable_intervals(double xmax, int max_intervals, double* h) {
  // h = 1/2 would err by 1.4e-5 (~200 FP32 ulps of W): beyond h = 1/4 the
  // kernels evaluate W exactly instead
  for (double hh : {0.125, 0.25}) {
    const double m = ceil(xmax / hh) + 1.0;
    if (m <= max_intervals) {
      *h = hh;
      return (int)m;
    }
  }
  *h = 0.0;
  return 0;
}

hipError_t launch_wtable(float4* coef, int M, double h, hipStream_t s) {
  if (M <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_wtable, dim3((M + 255) / 256), dim3(256), 0, s, coef, M, h);
  return hipGetLastError();
}

hipError_t launch_maxabs(const float* x, long long n, unsigned int* out, hipStream_t s) {
  hipError_t e = hipMemsetAsync(out, 0, sizeof(unsigned int), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_maxabs, dim3(grid1(n, 256, 8192)), dim3(256), 0, s, x, n, out);
  return hipGetLastError();
}

hipError_t launch_count_nonfinite(const float* x, long long n, int* flag, hipStream_t s) {
  hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_nonfinite, dim3(grid1(n, 256, 8192)), dim3(256), 0, s, x, n, flag);
  return hipGetLastError();
}

}  // namespace prc
