// Internal interface between the C-ABI host code and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace prc {

constexpr int kNHistory = 100000;  // NHISTORY, c++/src/allvars_d2z.h:1

// Reference parameters as the reference stores them (floats; BoxSize in
// Mpc/h as read from the parameter file, Rmax in Mpc/h, M_min in Msun).
struct SetupParams {
  float Omegam, Omegab, Omegal, h, ns, sigma8, w;
  float BoxSize, zInit, Rmin, Rmax, zeta, M_min;
  int N, Nscales;
};

// Scalars produced by the FP64 set-up kernels.
struct SetupScalars {
  float pnorm, sigma_min, erfizeta, DInit, DHigh, A;
};

// Per-radius constants read by the filter and barrier kernels.
struct RadiusConst {
  float R;      // Mpc
  float sigma;  // sigma(R), float as in the reference
  double A;     // (1+zHigh) D(zHigh)/deltac, widened from float
  double B;     // sqrt(2 (sigma_min^2 - sigma^2)) * erfizeta, double
};

// Synthetic field spectrum: BBKS with Sugiyama shape, k in h/Mpc.
struct GrfSpectrum {
  double gamma, ns, sigma8;
  double dk;     // 2 pi / BoxSize [h/Mpc]
  double vcell;  // (BoxSize/N)^3 [(Mpc/h)^3]
  double invN3;  // 1/N^3 of the inverse transform
};

hipError_t launch_twiddles(float2* tw, int L, hipStream_t s);
hipError_t launch_fwd_z(int Nh, const float* in, long long in_row_stride, float2* out, long long nrows,
                        float dinit, const float2* tw_half, const float2* tw_full, hipStream_t s);
hipError_t launch_inv_z(int Nh, float2* data, long long nrows, const float2* tw_half, const float2* tw_full,
                        hipStream_t s);
hipError_t launch_strided(int L, int dir, float2* data, long long sA, long long sO, long long nOuter, int ncols,
                          const float2* tw, hipStream_t s);
hipError_t launch_unpack(const float2* S, float2* D0, float2* DN, int N, hipStream_t s);
hipError_t launch_filter(const float2* S, const float2* D0, const float2* DN, float2* out, int N, float dkR,
                         float invN3, hipStream_t s);
hipError_t launch_barrier(const float* dR, float* zre, float* rbub, long long total, const RadiusConst* rc, int r,
                          hipStream_t s);
hipError_t launch_moments(const float* x, long long total, double2* partial, int nblk, hipStream_t s);
hipError_t launch_moments_finish(const double2* partial, int nblk, int nrad, double invN3, float* sigma_grid_out,
                                 int r0, hipStream_t s);
hipError_t launch_history(const float* zre, long long total, unsigned int* counts, double dz, hipStream_t s);
hipError_t launch_setup(const SetupParams& p, double* integ, float* Rr, double* Rint, SetupScalars* sc,
                        RadiusConst* rc, float* sigma_out, hipStream_t s);

// fused path
hipError_t launch_fx(int N, const float2* S, const float2* D0, const float2* DN, float2* Tbuf, float dk, float invN3,
                     const RadiusConst* rc, int r0, int nb, const float2* tw, hipStream_t s);
int fz_blocks(int Nh, long long nrows);
hipError_t launch_fz(int Nh, const float2* Tbuf, float* zre, float* rbub, long long nrows, const RadiusConst* rc,
                     int r0, int nb, double* partial, int nblk, const float2* tw_half, const float2* tw_full,
                     hipStream_t s);
hipError_t launch_count_nonfinite(const float* x, long long n, int* flag, hipStream_t s);
hipError_t launch_fill(float* p, long long n, float v, hipStream_t s);
hipError_t launch_sigma_finish(const double* partial, int nblk, int nrad, double invN3, float* sigma_grid_out,
                               hipStream_t s);

// synthetic input field
hipError_t launch_white_noise(float* out, long long n, unsigned long long seed, hipStream_t s);
hipError_t launch_grf_norm(const GrfSpectrum& sp, double* A, hipStream_t s);
hipError_t launch_colour(float2* S, const float2* D0, const float2* DN, int N, const GrfSpectrum& sp,
                         const double* A, hipStream_t s);

}  // namespace prc
