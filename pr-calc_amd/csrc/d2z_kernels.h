// Internal interface between the C-ABI host code and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace prc {

constexpr int kNHistory = 100000;  // NHISTORY, c++/src/allvars_d2z.h:1
constexpr int kNCHistory = 10000;  // NCHISTORY, c++/src/allvars_d2z.h:5
constexpr int kNXMfp = 1000;       // NX_MFP, c++/src/allvars_d2z.h:9
constexpr int kNMfp = 200;         // NMFP, c++/src/allvars_d2z.h:13
constexpr int kNSkewers = 20000;   // NSKEWERS, c++/src/allvars_d2z.h:17
// Start-cell draws per skewer before the call fails (the reference draws until
// it finds an ionized cell, without bound; at an ionized fraction of 1% the
// chance of 16384 misses is e^-164)
constexpr int kMaxStartDraws = 16384;

// Skewer walk set-up: grid size, skewers per ionized fraction, Philox key, cell
// size (BoxSize/N as float, widened) and the log-distance bins of the histogram
struct SkewerParams {
  int N, nskewers;
  unsigned int seed_lo, seed_hi;
  double cell;
  float lmin, lmax, dmfp;
};

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

constexpr int kFusedMaxRadii = 32;  // radii per fused batch (y and z passes)
constexpr int kFxMaxRadii = 16;     // radii per x-pass launch (a batch takes one or more)

// Strides of one tiled pass (complex elements); see pass_kernels.hip.
struct TileMap {
  long long in_sA, in_sO, out_sA, out_sO, sCB;
  int nOuter, nColBlocks;
};
enum class Axis { kYInPlace, kXTtoS, kXStoT };

// Cubic Hermite table of the top-hat window W(x) = 3 (sin x - x cos x) / x^3 on
// [0, M h): interval i holds (c0, c1, c2, c3) with W(x_i + t h) = c0 + t (c1 +
// t (c2 + t c3)), built from W and W' in FP64 at the nodes.  M = 0: no table
// (the kernels evaluate W directly).
struct WTable {
  const float4* coef;
  int M;
  float invh;  // 1 / h, a power of two
};
constexpr int kWTableMax = 3584;       // intervals (56 KB of LDS)
constexpr int kWTableSplitMax = 3184;  // the split x pass's room for it (its other LDS: 110 KB)
// intervals and spacing covering [0, xmax]: h = 1/8 if it fits in
// max_intervals, else 1/4; 0 intervals (exact W) if neither does.  Max |error|
// of the interpolant (W(0) = 1): 9.6e-8 at h = 1/8, 9.0e-7 at 1/4
int wtable_intervals(double xmax, int max_intervals, double* h);
hipError_t launch_wtable(float4* coef, int M, double h, hipStream_t s);
int block_width(int N);
TileMap tile_map(int N, Axis axis);

hipError_t launch_twiddles(float2* tw, int L, hipStream_t s);
// per-stage twiddle tables of the length-N and length-N/2 plans
int stage_twiddle_size(int N, bool half);
hipError_t launch_stage_twiddles(int N, float2* full, float2* half, hipStream_t s);
// matrix-core fused x pass (mixed path, N = 1024); tw_full_plan is the
// per-stage twiddle table of the length-N plan
bool fx_mfma_supported(int N, bool half);
hipError_t launch_fx_mfma(int N, const float2* S, const float2* D0, const float2* DN, void* Tbuf, float dk,
                          float invN3, const RadiusConst* rc, int r0, int nb, const float2* tw_full_plan,
                          const WTable& wt, const void* mats, hipStream_t s);
// matrix-core fused z pass (mixed path, N = 1024); tw_full = exp(2 pi i m / N)
bool fz_mfma_supported(int N, bool half);
int fz_mfma_blocks(bool rb);
hipError_t launch_fz_mfma(int N, const void* Tbuf, float* zre, float* rbub, const RadiusConst* rc, int r0, int nb,
                          double* partial, int nblk, float iscale, const float2* tw_full, const void* mats,
                          hipStream_t s);
// Compensated matrix-core x and z passes (mixed path, FP32 intermediates,
// N = 1024): split hi/lo FP16 operands of the radix-16 stages, FP32
// accumulation (mfma_split_kernels.hip)
bool split_supported(int N);
size_t split_matrix_bytes(int N);
hipError_t launch_split_matrices(void* out, hipStream_t s);
hipError_t launch_fx_split(int N, const float2* S, const float2* D0, const float2* DN, float2* T, float dk,
                           float invN3, const RadiusConst* rc, int r0, int nb, const float2* tw_full,
                           const WTable& wt, const void* mats, hipStream_t s);
int fz_split_blocks(bool rb);
// zscale: power of two <= 256 / max|delta| of the input field
hipError_t launch_fz_split(int N, const float2* T, float* zre, float* rbub, const RadiusConst* rc, int r0, int nb,
                           double* partial, int nblk, float zscale, const float2* tw_full, const void* mats,
                           hipStream_t s);
// FP16 operand matrices of the matrix-core passes' radix-16 stages, built
// once (0 bytes when neither pass runs for N and the storage type)
size_t mfma_matrix_bytes(int N, bool half);
hipError_t launch_mfma_matrices(void* out, hipStream_t s);
// per-stage twiddles of the fused x pass's plan
int fx_twiddle_size(int N);
hipError_t launch_fx_twiddles(int N, float2* out, hipStream_t s);
// pencil passes (pass_kernels.hip)
hipError_t launch_fwd_z(int N, const float* in, float2* out, float dinit, const float2* twh, const float2* twf,
                        hipStream_t s);
hipError_t launch_inv_z(int N, const float2* in, float* out, const float2* twh, const float2* twf, hipStream_t s);
hipError_t launch_tile(int N, int dir, const float2* in, float2* out, const TileMap& mp, const float2* tw,
                       hipStream_t s);
hipError_t launch_unpack(int N, const float2* S, float2* D0, float2* DN, hipStream_t s);
hipError_t launch_filter(int N, const float2* S, const float2* D0, const float2* DN, float2* out, float dkR,
                         float invN3, hipStream_t s);
// i k_a / |k|^2 * scale * S, component a = axis (0, 1, 2), k = n dk (the
// displacement field's spectrum, S layout as the filter's output)
hipError_t launch_disp(int N, const float2* S, const float2* D0, const float2* DN, float2* out, int axis, float dk,
                       float scale, hipStream_t s);
// binned power of the half spectrum (count[nk], sums[3][nk]: power * pscale,
// 1/k^2, k); clears both first (not for graph capture)
constexpr int kPkMaxBins = 512;
hipError_t launch_pk(int N, const float2* S, const float2* D0, const float2* DN, float dk, float lnkmin, float dlnk,
                     int nk, double pscale, unsigned long long* count, double* sums, hipStream_t s);
hipError_t launch_inv_z_h(int N, const void* in, float* out, float iscale, const float2* twh, const float2* twf,
                          hipStream_t s);
// in-place tiled pass over FP16 intermediates (mixed path), output scaled
hipError_t launch_tile_h(int N, void* data, const TileMap& mp, float oscale, const float2* tw, hipStream_t s);
// fused passes; Tbuf holds float2 (half = false) or __half2 (half = true)
hipError_t launch_fx(int N, const float2* S, const float2* D0, const float2* DN, void* Tbuf, bool half, float dk,
                     float invN3, const RadiusConst* rc, int r0, int nb, const float2* tw, const WTable& wt,
                     hipStream_t s);
// blocks of the fused z pass: the device's resident capacity for that kernel,
// at most kFzMaxBlocks and at most one per 256-thread row group.  fbar (FP32
// storage only): the barrier as one FP32 FMA instead of the double evaluation
constexpr int kFzMaxBlocks = 4096;
int fz_blocks(int N, bool half, bool rb, bool fbar);
hipError_t launch_fz(int N, const void* Tbuf, bool half, float* zre, float* rbub, const RadiusConst* rc, int r0,
                     int nb, double* partial, int nblk, float iscale, const float2* twh, const float2* twf, bool fbar,
                     hipStream_t s);
// ------------------------------------------------------ multi-box launches
// One box of a persistent multi-box launch (the buffers of its context).
// Boxes of N in {64, 128, 192, 256} run together: the blocks of a launch walk
// a list of work items (box, tile or row group) and every item runs the same
// device code as the single-box kernel of its N, so results are those of the
// box's own run.  The light passes take all boxes in one launch; the fused x
// and z passes take one launch per box size (a size class), so that each is
// compiled for one N and keeps that N's registers and occupancy.
struct MBox {
  const float* delta;
  float2* S;
  float2* D0;
  float2* DN;
  float2* T;
  float* zre;
  float* rbub;  // nullptr: bubble radii not recorded
  const RadiusConst* rc;
  double* part;  // [Nscales][kFzMaxBlocks] per-block sums of the z pass
  const float2* st_full;  // stage twiddles of the tile plan (x, y passes)
  const float2* st_half;  // stage twiddles of the half-length z plan
  const float2* tw_full;  // exp(2 pi i m / N), m < N
  const float2* st_fx;    // stage twiddles of the fused x plan
  const float4* wcoef;    // the box's W(x) table
  float* sigma_grid;      // [Nscales] output of the sums of squares
  int wM;
  float winvh, dk, invN3, dinit;
  int N, Nscales;
  int zgrid;  // grid of the z pass of this box's size class: the stride of part
};
constexpr int kMBoxMax = 256;  // boxes per multi-box launch
bool mbox_supported(int N);
// passes of a multi-box step; `first` holds nbox + 1 item offsets per pass
enum class MBPass { kFill, kFwdZ, kFwdY, kFwdX, kUnpack, kFx, kInvY, kFz, kSigma };
long long mbox_items(MBPass pass, int N, int nb);
// resident capacity of the pass's kernel (persistent grid); N: the size class
// of the x and z passes (ignored by the others)
int mbox_blocks(MBPass pass, bool rb, int N);
// boxes[0..nbox) with item offsets first[0..nbox]; r0, nb: the batch's radii
// (x pass: this launch's radii); toff: the x pass's first intermediate plane
// within the batch; maxM: the largest W table of the boxes (x pass LDS); N:
// the size class of an x or z pass launch (every box of it has that N)
hipError_t launch_mbox(MBPass pass, const MBox* boxes, const long long* first, int nbox, long long nitems, int r0,
                       int nb, int toff, int maxM, int nblk, int N, bool rb, hipStream_t s);

// element-wise and set-up kernels (d2z_kernels.hip)
hipError_t launch_barrier(const float* dR, float* zre, float* rbub, long long total, const RadiusConst* rc, int r,
                          hipStream_t s);
hipError_t launch_moments(const float* x, long long total, double2* partial, int nblk, hipStream_t s);
hipError_t launch_moments_finish(const double2* partial, int nblk, int nrad, double invN3, float* sigma_grid_out,
                                 int r0, hipStream_t s);
hipError_t launch_history(const float* zre, long long total, unsigned int* counts, double dz, int* range,
                          hipStream_t s);
// history correction (bubble_kernels.hip)
hipError_t launch_skewers(const float* zre, const float* zcur, const SkewerParams& p, double* sums,
                          unsigned int* counts, int* fail, hipStream_t s);
hipError_t launch_zvz(const float* history, const float* chist, float dzch, float* zvz, hipStream_t s);
hipError_t launch_remap(float* zre, long long n, const float* zvz, double dz, hipStream_t s);
hipError_t launch_setup(const SetupParams& p, double* integ, float* Rr, double* Rint, SetupScalars* sc,
                        RadiusConst* rc, float* sigma_out, hipStream_t s);

hipError_t launch_count_nonfinite(const float* x, long long n, int* flag, hipStream_t s);
// *out = bits of max |x| (not for graph capture: clears with a memset)
hipError_t launch_maxabs(const float* x, long long n, unsigned int* out, hipStream_t s);
hipError_t launch_fill(float* p, long long n, float v, hipStream_t s);
hipError_t launch_sigma_finish(const double* partial, int nblk, int nrad, double invN3, float* sigma_grid_out,
                               hipStream_t s);

// synthetic input field
hipError_t launch_white_noise(float* out, long long n, unsigned long long seed, hipStream_t s);
hipError_t launch_grf_norm(const GrfSpectrum& sp, double* A, hipStream_t s);
hipError_t launch_colour(int N, float2* S, const float2* D0, const float2* DN, const GrfSpectrum& sp,
                         const double* A, hipStream_t s);

}  // namespace prc
