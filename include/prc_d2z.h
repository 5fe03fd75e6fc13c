/* prc_d2z.h -- C ABI of the MI355X delta2zreion library (libprc_d2z.so).
 *
 * The library replaces two seams of pr-calc's d2z program
 * (marcelo-alvarez/pr-calc):
 *   1. the FFTW3-MPI set-up and forward transform in c++/src/main_d2z.C:32-48
 *      (fftwf_mpi_local_size_3d, AllocateArrays, plan creation) and :52-67
 *      (delta /= DInit, fftwf_execute(plan));
 *   2. void delta2zreion() (c++/src/proto_d2z.h:27, c++/src/delta2zreion_d2z.C:8-216):
 *      sigma normalisation, erfc^-1(1/zeta), sigma_min, and for every radius the
 *      top-hat filter, c2r, 1/N^3, sigma_grid, sigma(R), barrier and running max.
 * The host keeps the parameter file, command line, ICs input, history and
 * zreion output.  One context owns the device buffers of one box; the whole
 * grid lives on one GPU (no MPI, no slab decomposition).
 *
 * Conventions: every call returns PRC_D2Z_OK (0) or a negative status and never
 * prints or exits; prc_d2z_last_error() gives the message.  Host buffers are
 * caller-owned.  Calls on one context are synchronous unless named _async and
 * must not overlap; different contexts are independent.
 */
#ifndef PRC_D2Z_H
#define PRC_D2Z_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRC_D2Z_ABI_VERSION 1
#define PRC_D2Z_NHISTORY 100000 /* NHISTORY, c++/src/allvars_d2z.h:1 */
#define PRC_D2Z_NCHISTORY 10000 /* NCHISTORY, c++/src/allvars_d2z.h:5 */
#define PRC_D2Z_NX_MFP 1000     /* NX_MFP, c++/src/allvars_d2z.h:9 */
#define PRC_D2Z_NMFP 200        /* NMFP, c++/src/allvars_d2z.h:13 */
#define PRC_D2Z_NSKEWERS 20000  /* NSKEWERS, c++/src/allvars_d2z.h:17 */

typedef struct prc_d2z_ctx prc_d2z_ctx;

enum prc_d2z_status {
  PRC_D2Z_OK = 0,
  PRC_D2Z_ERR_BAD_ARG = -1,
  PRC_D2Z_ERR_UNSUPPORTED_N = -2,
  PRC_D2Z_ERR_OOM = -3,
  PRC_D2Z_ERR_HIP = -4,
  PRC_D2Z_ERR_NAN_INPUT = -5,
  PRC_D2Z_ERR_STATE = -6,
  PRC_D2Z_ERR_NO_DEVICE = -7
};

/* Execution paths over the same arithmetic contract. */
enum prc_d2z_mode {
  PRC_D2Z_PLAIN_FP32 = 0, /* separate kernels: filter, 3 c2r passes, barrier, moments */
  PRC_D2Z_FUSED_FP32 = 1, /* filter fused into the first inverse pass, barrier/max into the last,
                             radii_per_launch radii per launch off one resident spectrum */
  PRC_D2Z_FUSED_MIXED = 2, /* fused; at N = 1024 the x and z passes' radix-16 stages run on the matrix
                             cores with split hi/lo FP16 operands and FP32 accumulation (FP32-grade
                             transforms), FP32 intermediates; other N as FUSED_FP32 */
  PRC_D2Z_FUSED_HALF = 3   /* fused, with FP16 storage of the inter-pass intermediates (and, at N = 1024,
                             plain FP16 matrix-core radix-16 stages): narrower than FP32 */
};

/* Real-space field layouts of the reference: dense N^3 floats (ICs and .zreion
 * files, c++/src/io_d2z.C:164-246) or the in-place FFTW layout float[N][N][N+2]
 * (c++/src/allocate_d2z.C:27-37). */
enum prc_d2z_layout { PRC_D2Z_DENSE = 0, PRC_D2Z_PADDED = 1 };

typedef struct {
  /* parameter-file tags, c++/src/io_d2z.C:33-87 (struct Parameter,
   * c++/src/allvars_d2z.h:58-64); BoxSize in Mpc/h as in the file */
  float Omegam, Omegab, Omegal, h, ns, sigma8, w;
  float BoxSize, zInit, Rmin;
  int N, Nscales;
  int RecordBubbleSizes;
  /* command-line values (struct clParameter, c++/src/allvars_d2z.h:66-70;
   * defaults c++/src/commandline_d2z.C:47-52): Rmax in Mpc/h, M_min in Msun */
  float Rmax, zeta, M_min;
  /* execution options (no counterpart in the reference) */
  int mode;             /* enum prc_d2z_mode */
  int radii_per_launch; /* fused modes: radii per batch, 1..32; 0 = default */
  int device;           /* HIP device ordinal */
} prc_d2z_params;

/* Derived scalars (globals DInit, sigma_min, erfizeta, zHigh, DHigh of
 * c++/src/allvars_d2z.h:76 plus pnorm and the barrier slope A). */
typedef struct {
  float DInit, DHigh, A, pnorm, erfizeta, sigma_min;
} prc_d2z_scalars;

/* One row of the per-radius table printed by delta2zreion_d2z.C:102-103. */
typedef struct {
  float R_mpch;     /* R * h */
  float sigma;      /* sigma(R) from the EH98 no-wiggle spectrum */
  float sigma_grid; /* rms of the smoothed grid */
} prc_d2z_radius_row;

/* The same row with the barrier constants of d2z() (delta2zreion_d2z.C:206-216):
 * z = A (delta_R + B) - 1, A = (1+zHigh) D(zHigh)/deltac, B = sqrt(2 (sigma_min^2 -
 * sigma^2)) erfizeta; B is NaN where sigma(R) > sigma_min, as the reference's
 * square root (such radii never raise zreion). */
typedef struct {
  float R_mpch, sigma, sigma_grid;
  double A, B;
} prc_d2z_radius_diag;

/* Kernel timing of the last prc_d2z_run, from HIP events recorded on the
 * context's stream around every launch. */
#define PRC_D2Z_MAX_KERNELS 16
typedef struct {
  double total_ms; /* forward transform through the last radius */
  int nkernels;
  char name[PRC_D2Z_MAX_KERNELS][32];
  double ms[PRC_D2Z_MAX_KERNELS]; /* summed over all launches of that kernel */
  int launches[PRC_D2Z_MAX_KERNELS];
} prc_d2z_timing;

/* Fills reference defaults (commandline_d2z.C:47-52: Rmax 256, zeta 10,
 * M_min 1e8) and a 128^3 / 200 Mpc/h / 32-radius box. */
void prc_d2z_default_params(prc_d2z_params* p);

/* fftwf_mpi_local_size_3d + AllocateArrays + plan creation + the scalar
 * set-up of delta2zreion (normalize, inverse_erfc, sigma_min, sigma(R)),
 * computed on the GPU in FP64.  N must be 2^a 3^b with N/2 supported
 * (16, 32, 64, 128, 192, 256, 512, 1024, 2048). */
int prc_d2z_create(const prc_d2z_params* p, prc_d2z_ctx** out);
void prc_d2z_destroy(prc_d2z_ctx* ctx);
/* Message of the last failed call on ctx (ctx == NULL: last failed create on
 * this thread). */
const char* prc_d2z_last_error(const prc_d2z_ctx* ctx);

/* ReadDeltaFile's destination (io_d2z.C:164-207): copy the real-space density
 * contrast (N^3 cells, x slowest) to the device.  Rejects NaN/Inf. */
int prc_d2z_set_delta(prc_d2z_ctx* ctx, const float* delta, int layout);
/* Same from device memory already on the context's device. */
int prc_d2z_set_delta_device(prc_d2z_ctx* ctx, const float* d_delta, int layout);

/* Synthetic linear density field generated on the device into the context's
 * input buffer (replaces the ics program's output, c++/src/grf_ics.C:8-112):
 * Philox4x32-10 normals keyed by (seed, cell), coloured by
 * sqrt(P(k)/V_cell) with P = A k^ns T_BBKS(k)^2 (Sugiyama shape) normalised to
 * sigma8 from the context's cosmology and BoxSize.  white_noise_only = 1 stops
 * after the normals. */
int prc_d2z_generate_delta(prc_d2z_ctx* ctx, uint64_t seed, int white_noise_only);
/* Copy the context's current real-space input field back to the host. */
int prc_d2z_get_delta(prc_d2z_ctx* ctx, float* out, int layout);

/* main_d2z.C:57-71: delta /= DInit, forward r2c, then delta2zreion() for all
 * radii.  zreion starts from 0 (allocate_d2z.C:37) on every call. */
int prc_d2z_run(prc_d2z_ctx* ctx);
int prc_d2z_run_async(prc_d2z_ctx* ctx);
int prc_d2z_sync(prc_d2z_ctx* ctx);

/* Results (WriteZreionFile's source, io_d2z.C:209-246). */
int prc_d2z_get_zreion(prc_d2z_ctx* ctx, float* out, int layout);
/* rbubble (RecordBubbleSizes = 1): R in Mpc of the radius that set zreion; 0
 * where zreion stayed 0. */
int prc_d2z_get_rbubble(prc_d2z_ctx* ctx, float* out, int layout);
int prc_d2z_get_scalars(prc_d2z_ctx* ctx, prc_d2z_scalars* out);
int prc_d2z_get_radius_table(prc_d2z_ctx* ctx, prc_d2z_radius_row* rows, int n);
int prc_d2z_get_radius_diag(prc_d2z_ctx* ctx, prc_d2z_radius_diag* rows, int n);
/* Pins the scalar set-up of delta2zreion (delta2zreion_d2z.C:17-26,97-100) to
 * host values, e.g. the reference's own GSL results: sigma[0] = sigma_min,
 * sigma[1 + r] = sigma(R_r) (r < Nscales), *erfizeta = inverse_erfc(1/zeta);
 * either pointer may be NULL to keep the computed values.  B_R is rebuilt from
 * them; the next prc_d2z_run uses them. */
int prc_d2z_set_scalar_overrides(prc_d2z_ctx* ctx, const float* sigma /* [Nscales + 1] */, const float* erfizeta);
/* Device pointer of the dense zreion field (valid until destroy). */
int prc_d2z_zreion_device(prc_d2z_ctx* ctx, const float** d_zreion);

/* GetHistory (c++/src/bubblehistory_d2z.C:5-41) on the GPU: integer counts per
 * history bin, bin = (int)((zreion - 49.9)/dzhistory). */
int prc_d2z_history(prc_d2z_ctx* ctx, uint32_t* counts /* [PRC_D2Z_NHISTORY] */);

/* GetHistory (as prc_d2z_history) of a zreion field the caller holds in device
 * memory of the current device (ncells floats, any order, 16-byte aligned), without a context
 * and without any allocation: it calls no hipMalloc, hipFree or allocating
 * library routine.  No counterpart in the reference: it is for callers that
 * cannot allocate (serving loops, harnesses that forbid it).  Its device
 * scratch is a static array of NHISTORY + 64 32-bit words per GPU (~400 KB,
 * held until the process ends); ncells is limited to 2^32 - 1 (32-bit counts).
 * Runs on the null stream and returns when counts[PRC_D2Z_NHISTORY] is filled.
 * Calls that overlap on one GPU race on that scratch.  Errors: BAD_ARG (null,
 * ncells out of range, d_zreion not device memory of the current device),
 * NO_DEVICE, HIP; the message is prc_d2z_last_error(NULL). */
int prc_d2z_history_noalloc(const float* d_zreion, int64_t ncells, uint32_t* counts);

/* History correction of the written zreion field (main_d2z.C:78-90), in the
 * reference's order: GetHistory (above, history = cumulative counts / N^3 as
 * float), GetBubbleMFP, GetCHistory, CorrectHistory.
 *
 * prc_d2z_bubble_mfp -- GetBubbleMFP (c++/src/bubblehistory_d2z.C:82-242) on
 * the GPU for one rank: for each ionized fraction x_i = 0.01 + i dx (i <
 * PRC_D2Z_NX_MFP, dx = 0.989/999) and zcur_i where the float history
 * [PRC_D2Z_NHISTORY] crosses x_i, nskewers skewers (0 = PRC_D2Z_NSKEWERS) start
 * at random cells with zreion > zcur_i (j, k < N - 2), walk along +y and +z to
 * the first cell with zreion < zcur_i, and record a distance log-uniform in
 * [m - 1, m] cells.  Random numbers: Philox4x32-10 keyed by seed, counter
 * (skewer, i, draw) -- the reference's rand() sequence is not reproduced.
 * mfp[i] = mean distance in Mpc (the cell is BoxSize/h/N: the d2z program
 * divides the parameter file's BoxSize by h first, main_d2z.C:29), mfp_x[i] =
 * x_i, dfdmfp[i * PRC_D2Z_NMFP + b] = walks in ln-distance bin b between
 * ln(cell) and ln(BoxSize/h) (mfp_x, dfdmfp nullable).  A skewer that finds
 * no start cell in 16384 draws fails the call (the reference loops forever). */
int prc_d2z_bubble_mfp(prc_d2z_ctx* ctx, const float* history, int nskewers, uint64_t seed, float* mfp,
                       float* mfp_x, float* dfdmfp);
/* prc_d2z_corrected_history -- GetCHistory (c++/src/bubblehistory_d2z.C:44-80)
 * on the host: chistory[PRC_D2Z_NCHISTORY] from the collapsed-fraction table
 * fcoll(z) (the file fcoll_table.tab: float min, float max, int n, float
 * table[n]; Float2FloatTable, c++/src/tables.C:9-38), the mean free paths of
 * prc_d2z_bubble_mfp and the context's Rmax and zeta. */
int prc_d2z_corrected_history(prc_d2z_ctx* ctx, const float* fcoll, int nfcoll, float fmin, float fmax,
                              const float* mfp, float* chistory);
/* prc_d2z_correct_zreion -- CorrectHistory (c++/src/bubblehistory_d2z.C:270-343)
 * on the GPU: every cell's zreion becomes the redshift at which the corrected
 * history reaches the history's value in the cell's bin; bins are clamped to
 * the table (the reference reads out of bounds for zreion < 0.01 or > 49.9).
 * Replaces the context's zreion (prc_d2z_get_zreion returns the corrected
 * field). */
int prc_d2z_correct_zreion(prc_d2z_ctx* ctx, const float* history, const float* chistory);

/* Smoothed field delta_R of radius r (filter + c2r + 1/N^3), dense layout,
 * through the context's transform path. */
int prc_d2z_smoothed_field(prc_d2z_ctx* ctx, int r, float* out);

/* Linear displacement field of the context's density field, component axis
 * (0 = x, 1 = y, 2 = z): psi_a = IFFT(i k_a / k^2 FFT(delta)) / N^3 with
 * k = 2 pi n / BoxSize (BoxSize as in the parameters), k = 0 and the
 * component's own Nyquist plane zeroed -- Delta2Velocity's xdv_a
 * (c++/src/delta2velocity_map.C:8-97, its two N^-1.5 normalisations
 * included), which the map programs scale into velocities.  FP32 transforms;
 * out: N^3 floats in `layout` (enum prc_d2z_layout). */
int prc_d2z_displacement(prc_d2z_ctx* ctx, int axis, float* out, int layout);

/* Binned power spectrum of the context's density field, PowerSpectrum
 * (c++/src/powerspectrum_lmb.C:8-139, transform = 1): modes of the half
 * spectrum (kz = 0..N/2, each once) of FFT(delta) N^-1.5, k = 2 pi n /
 * BoxSize in the reference's float arithmetic, bin = (int)((log k - log kmin)
 * / dlnk) with dlnk = (log kmax - log kmin) / nk.  Per bin (nk <= 512): ps =
 * (2 pi)^3 <|delta_k|^2> / (dk N)^3, ook2 = <1/k^2>, kmean = <k>, nmodes =
 * count; empty bins give NaN as the reference's divisions by zero. */
int prc_d2z_power_spectrum(prc_d2z_ctx* ctx, int nk, float kmin, float kmax, double* ps, double* ook2,
                           double* kmean, int64_t* nmodes);

/* Launch shape of the fused modes: radii per fused batch (1..max_radii, bounded
 * by the intermediate buffers allocated at create) and the grid of the fused
 * z pass (1..4096 blocks). */
int prc_d2z_set_launch_config(prc_d2z_ctx* ctx, int radii_per_launch, int z_blocks);
int prc_d2z_get_launch_config(const prc_d2z_ctx* ctx, int* radii_per_launch, int* z_blocks, int* max_radii);

/* FFT plans of the fused x and z passes: the vector radix plans (radix 2, 3,
 * 4, 8, 16 stages on the vector ALUs) or the matrix-core plans (radix-16
 * stages on the matrix cores: split hi/lo FP16 operands in the mixed mode at
 * N = 1024; plain FP16 in the half mode, x pass at N = 1024 and 2048, z pass
 * at N = 1024).  Default: matrix-core where the mode and N have them; asking
 * for one that does not exist returns UNSUPPORTED_N.  x_radii_per_launch:
 * radii per x-pass launch, 1..16. */
enum prc_d2z_plan { PRC_D2Z_PLAN_VECTOR = 0, PRC_D2Z_PLAN_MATRIX = 1 };
int prc_d2z_set_plan(prc_d2z_ctx* ctx, int x_plan, int z_plan, int x_radii_per_launch);
int prc_d2z_get_plan(const prc_d2z_ctx* ctx, int* x_plan, int* z_plan, int* x_radii_per_launch);

typedef struct {
  int radii_per_launch;
  int z_blocks;
  double ms_per_radius;
  int x_plan, z_plan;     /* enum prc_d2z_plan */
  int x_radii_per_launch;
} prc_d2z_tuning;
/* Autotuner: times the fused batches of the first radii for each candidate
 * (x and z FFT plans, radii per x launch, radii per batch, z grid) on the
 * current field, keeps and returns the fastest.  The host persists the choice
 * (see prc_d2z.tuning in the Python binding). */
int prc_d2z_autotune(prc_d2z_ctx* ctx, prc_d2z_tuning* out);

/* Runs several contexts (boxes of any size on the same device) at once: every
 * context's captured run is submitted on its own stream, then all are awaited. */
int prc_d2z_run_many(prc_d2z_ctx* const* ctxs, int n);

/* Multi-box group: boxes of N = 64, 128, 192 or 256 (mode FUSED_FP32, or
 * FUSED_MIXED, whose kernels are the FP32 ones at these N), the same Nscales
 * and radii per batch, run together -- the boxes are spread over four chains
 * (streams) balanced by cells; on a chain, the boxes of one size go through
 * one launch per pass that walks a work list of (box, tile) items with each
 * box's own plan; the chains run side by side and the whole step is one
 * captured graph.  The boxes keep their contexts (create, set_delta or
 * generate_delta, get_zreion as usual); results equal each box's prc_d2z_run
 * bit for bit (sigma_grid may differ in the last bits: its partial sums are
 * grouped differently).  The group holds the contexts (and the buffers and
 * launch shapes they had at create): destroy it before any of them, and
 * create a new group after changing a box's launch configuration.  No
 * counterpart in the reference, which runs one box per process. */
typedef struct prc_d2z_group prc_d2z_group;
int prc_d2z_group_create(prc_d2z_ctx* const* ctxs, int n, prc_d2z_group** out);
int prc_d2z_group_run(prc_d2z_group* grp);
void prc_d2z_group_destroy(prc_d2z_group* grp);
/* Message of the last failed call on grp (grp == NULL: last failed create). */
const char* prc_d2z_group_last_error(const prc_d2z_group* grp);

/* Replay prc_d2z_run as one captured hipGraph (on by default; the graph is
 * captured on the first run of the context).  Timing mode runs eagerly. */
int prc_d2z_set_graph(prc_d2z_ctx* ctx, int enable);

/* Per-kernel HIP-event timing of prc_d2z_run (off by default). */
int prc_d2z_set_timing(prc_d2z_ctx* ctx, int enable);
int prc_d2z_get_timing(prc_d2z_ctx* ctx, prc_d2z_timing* out);

/* Device buffer bytes this context holds. */
size_t prc_d2z_device_bytes(const prc_d2z_ctx* ctx);
int prc_d2z_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PRC_D2Z_H */
