// C-ABI implementation of libprc_d2z: context, device buffers, the radius
// loop orchestration and per-kernel event timing.
//
// Reference call sequence replaced (marcelo-alvarez/pr-calc):
//   main_d2z.C:32-48  fftwf_mpi_local_size_3d, AllocateArrays, plans  -> prc_d2z_create
//   main_d2z.C:52-59  ReadDeltaFile destination, delta /= DInit       -> prc_d2z_set_delta + run
//   main_d2z.C:67     fftwf_execute(plan)                             -> run: forward passes
//   main_d2z.C:71     delta2zreion()                                  -> run: radius loop
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/prc_d2z.h"
#include "d2z_device.h"
#include "d2z_kernels.h"

using namespace prc;

namespace {

thread_local std::string g_create_error;

struct KernelTimer {
  // event pairs recorded around launches while timing is on
  struct Rec {
    int cat;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::string> cats;
  hipEvent_t t0 = nullptr, t1 = nullptr;

  hipEvent_t get() {
    if (used == pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[used++];
  }
  int cat(const char* name) {
    for (size_t i = 0; i < cats.size(); ++i)
      if (cats[i] == name) return (int)i;
    cats.emplace_back(name);
    return (int)cats.size() - 1;
  }
  void reset() {
    recs.clear();
    used = 0;
  }
  void destroy() {
    for (auto e : pool) (void)hipEventDestroy(e);
    pool.clear();
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    t0 = t1 = nullptr;
  }
};

constexpr int kDefaultRadiiPerLaunch = kFusedMaxRadii;
constexpr int kMaxZBlocks = kFzMaxBlocks;  // largest grid of the fused z pass
constexpr int kMomentBlocks = 1024;

bool supported_n(int N) {
  switch (N) {
    case 16: case 32: case 64: case 128: case 192: case 256: case 512: case 1024: case 2048:
      return true;
    default:
      return false;
  }
}

// FP16 storage of the inter-pass intermediates (the half mode only)
bool half_storage(const prc_d2z_params& p) { return p.mode == PRC_D2Z_FUSED_HALF; }
// the compensated split-FP16 matrix-core x and z passes (mixed mode, FP32 storage)
bool use_split(const prc_d2z_params& p) { return p.mode == PRC_D2Z_FUSED_MIXED && split_supported(p.N); }
// a matrix-core plan exists for the mode, N and pass: the split passes (mixed,
// N = 1024) or the FP16 ones (half: x at N = 1024 and 2048, z at 1024);
// otherwise the pass runs its vector plan
bool has_matrix_x(const prc_d2z_params& p) { return use_split(p) || (half_storage(p) && fx_mfma_supported(p.N, true)); }
bool has_matrix_z(const prc_d2z_params& p) { return use_split(p) || (half_storage(p) && fz_mfma_supported(p.N, true)); }
// the mixed mode's vector z pass evaluates the barrier as one FP32 FMA, as its
// matrix-core z pass does; the fused FP32 mode keeps the reference's double one
bool fp32_barrier(const prc_d2z_params& p) { return p.mode == PRC_D2Z_FUSED_MIXED; }

}  // namespace

struct prc_d2z_ctx {
  prc_d2z_params p{};
  int N = 0, Nh = 0;
  long long ncell = 0, nspec = 0;  // N^3 reals, N*N*N/2 complex
  hipStream_t stream = nullptr;
  float* d_delta = nullptr;
  float2* d_S = nullptr;
  float2* d_D0 = nullptr;
  float2* d_DN = nullptr;
  float2* d_T = nullptr;
  int nbatch = 1;      // radii per fused batch
  int nbatch_cap = 1;  // most radii per batch the intermediate buffers hold
  size_t tbytes = 0;  // bytes of the intermediate buffer region
  float* d_zre = nullptr;
  float* d_rb = nullptr;
  float2* d_tw_full = nullptr;  // exp(+2 pi i k/N), k < N: real-transform glue
  float2* d_st_full = nullptr;  // per-stage twiddles of the length-N plan
  float2* d_st_half = nullptr;  // per-stage twiddles of the length-N/2 plan
  float2* d_st_fx = nullptr;    // per-stage twiddles of the fused x pass's plan
  float4* d_wcoef = nullptr;    // W(x) interpolation table of the fused x pass
  void* d_mats = nullptr;       // operand matrices of the matrix-core passes
  WTable wt{nullptr, 0, 0.0f};
  double2* d_mom = nullptr;   // plain: [Nscales][kMomentBlocks]
  double* d_part = nullptr;   // fused: [Nscales][kFzMaxBlocks], strided by the z pass grid
  int nblk_fz = 0;
  // FFT plans of the fused x and z passes (PRC_D2Z_PLAN_*) and radii per x launch
  int xplan = PRC_D2Z_PLAN_VECTOR, zplan = PRC_D2Z_PLAN_VECTOR;
  int fx_nb = kFxMaxRadii;
  float* d_sigma_grid = nullptr;
  float* d_sigma = nullptr;
  RadiusConst* d_rc = nullptr;
  SetupScalars* d_sc = nullptr;
  double* d_integ = nullptr;
  double* d_grfA = nullptr;
  double* d_Rint = nullptr;
  float* d_R = nullptr;
  unsigned* d_hist = nullptr;
  int* d_hrange = nullptr;  // lowest and highest occupied history bin
  double* d_pk = nullptr;  // power spectrum bins: counts (uint64) then 3 sums, kPkMaxBins each
  int* d_flag = nullptr;
  unsigned* d_maxabs = nullptr;  // max |delta| of the input field (float bits)
  // history correction: skewer sums and counts per ionized fraction, zcur per
  // fraction, the history in, the corrected history and the per-bin table
  double* d_mfp_sum = nullptr;
  unsigned* d_mfp_cnt = nullptr;
  float* d_zcur = nullptr;
  float* d_hin = nullptr;
  float* d_chist = nullptr;
  float* d_zvz = nullptr;
  SetupScalars sc{};
  float zscale = 1.0f;  // split z pass: input scale, a power of two <= 256 DInit / max|delta|
  std::vector<RadiusConst> rc;
  bool have_delta = false, have_spectrum = false, have_result = false;
  std::string err;
  size_t bytes = 0;
  bool timing = false;
  KernelTimer timer;
  bool use_graph = true;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<void*> allocs;

  int fail(int code, const std::string& msg) {
    err = msg;
    return code;
  }
  int hip_fail(hipError_t e, const char* what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return PRC_D2Z_ERR_HIP;
  }
  template <class T>
  hipError_t alloc(T** ptr, size_t bytes_) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes_);
    if (e == hipSuccess) {
      allocs.push_back(q);
      bytes += bytes_;
      *ptr = (T*)q;
    }
    return e;
  }
};

namespace {

// Launch wrapper that records timing events when enabled.
template <class F>
hipError_t timed(prc_d2z_ctx* c, const char* name, F&& f) {
  if (!c->timing) return f();
  KernelTimer& t = c->timer;
  hipEvent_t a = t.get(), b = t.get();
  if (!a || !b) return hipErrorOutOfMemory;
  (void)hipEventRecord(a, c->stream);
  hipError_t e = f();
  (void)hipEventRecord(b, c->stream);
  t.recs.push_back({t.cat(name), a, b});
  return e;
}

#define PRC_CHECK(ctx, expr, what)                       \
  do {                                                   \
    hipError_t e_ = (expr);                              \
    if (e_ != hipSuccess) return (ctx)->hip_fail(e_, what); \
  } while (0)

int forward(prc_d2z_ctx* c, float dinit) {
  const int N = c->N;
  hipStream_t s = c->stream;
  float2* T0 = c->d_T;
  const TileMap ym = tile_map(N, Axis::kYInPlace), xm = tile_map(N, Axis::kXTtoS);
  PRC_CHECK(c, timed(c, "fwd_z_r2c", [&] {
              return launch_fwd_z(N, c->d_delta, T0, dinit, c->d_st_half, c->d_tw_full, s);
            }),
            "forward z pass");
  PRC_CHECK(c, timed(c, "fwd_y", [&] { return launch_tile(N, -1, T0, T0, ym, c->d_st_full, s); }),
            "forward y pass");
  PRC_CHECK(c, timed(c, "fwd_x", [&] { return launch_tile(N, -1, T0, c->d_S, xm, c->d_st_full, s); }),
            "forward x pass");
  PRC_CHECK(c, timed(c, "unpack", [&] { return launch_unpack(N, c->d_S, c->d_D0, c->d_DN, s); }), "unpack");
  c->have_spectrum = true;
  return PRC_D2Z_OK;
}

// grid of the fused z pass: the resident capacity of the kernel the plan uses
int default_fz_blocks(const prc_d2z_params& p, int zplan) {
  const bool half = half_storage(p), rb = p.RecordBubbleSizes != 0;
  if (zplan == PRC_D2Z_PLAN_MATRIX && has_matrix_z(p)) return use_split(p) ? fz_split_blocks(rb) : fz_mfma_blocks(rb);
  return fz_blocks(p.N, half, rb, fp32_barrier(p));
}
bool matrix_x(const prc_d2z_ctx* c) { return c->xplan == PRC_D2Z_PLAN_MATRIX && has_matrix_x(c->p); }
bool matrix_z(const prc_d2z_ctx* c) { return c->zplan == PRC_D2Z_PLAN_MATRIX && has_matrix_z(c->p); }

// BoxSize in Mpc as the d2z program holds it after Parameters.BoxSize /= h
// (main_d2z.C:29, float division)
float box_mpc(const prc_d2z_params& p) { return p.BoxSize / p.h; }

float host_dk(const prc_d2z_ctx* c) {
  // float dk = 2.*M_PI/BoxSize with BoxSize already divided by h (delta2zreion_d2z.C:41)
  return (float)(2. * M_PI / (double)box_mpc(c->p));
}

// Plain path, one radius: separate filter, three inverse passes, moments and
// barrier.  T1 holds the filtered spectrum (S layout) and then the real field,
// T2 the intermediate (T layout).
int plain_radius(prc_d2z_ctx* c, int r, bool barrier) {
  const int N = c->N;
  hipStream_t s = c->stream;
  const float dkR = host_dk(c) * c->rc[r].R;
  const float invN3 = (float)(1.0 / ((double)N * N * N));
  float2* T1 = c->d_T;
  float2* T2 = c->d_T + c->nspec;
  const TileMap xm = tile_map(N, Axis::kXStoT), ym = tile_map(N, Axis::kYInPlace);
  PRC_CHECK(c, timed(c, "filter", [&] {
              return launch_filter(N, c->d_S, c->d_D0, c->d_DN, T1, dkR, invN3, s);
            }),
            "filter");
  PRC_CHECK(c, timed(c, "inv_x", [&] { return launch_tile(N, +1, T1, T2, xm, c->d_st_full, s); }),
            "inverse x pass");
  PRC_CHECK(c, timed(c, "inv_y", [&] { return launch_tile(N, +1, T2, T2, ym, c->d_st_full, s); }),
            "inverse y pass");
  float* dR = reinterpret_cast<float*>(T1);
  PRC_CHECK(c, timed(c, "inv_z_c2r", [&] {
              return launch_inv_z(N, T2, dR, c->d_st_half, c->d_tw_full, s);
            }),
            "inverse z pass");
  if (!barrier) return PRC_D2Z_OK;
  PRC_CHECK(c, timed(c, "moments", [&] {
              return launch_moments(dR, c->ncell, c->d_mom + (long long)r * kMomentBlocks, kMomentBlocks, s);
            }),
            "moments");
  PRC_CHECK(c, timed(c, "barrier", [&] {
              return launch_barrier(dR, c->d_zre, c->p.RecordBubbleSizes ? c->d_rb : nullptr, c->ncell, c->d_rc, r,
                                    s);
            }),
            "barrier");
  return PRC_D2Z_OK;
}

// Scales of the mixed path (powers of two, so rescaling is exact): the fused
// x pass multiplies by s1 = 2^round(log2 N), the y pass by s2 =
// 2^-round(log2 sqrt N), the z pass undoes both; intermediates then stay near
// the magnitude of delta_R, far inside the FP16 range.
struct MixedScales {
  float s1, s2, inv;
};
MixedScales mixed_scales(int N) {
  const int e1 = (int)lround(log2((double)N));
  const int e2 = (int)lround(0.5 * log2((double)N));
  return {ldexpf(1.0f, e1), ldexpf(1.0f, -e2), ldexpf(1.0f, e2 - e1)};
}

// The x pass of one batch (filter + x inverse) into the intermediates T, on s:
// launches of at most kFxMaxRadii radii (a launch's blocks write one plane per
// radius; 32 planes per block measured slower), each into its planes of T
int fused_x(prc_d2z_ctx* c, int r0, int nb, float2* T, hipStream_t s) {
  const int N = c->N;
  const bool half = half_storage(c->p);
  const MixedScales ms = half ? mixed_scales(N) : MixedScales{1.0f, 1.0f, 1.0f};
  const float invN3 = (float)(1.0 / ((double)N * N * N)) * ms.s1;
  const size_t plane_bytes = (half ? sizeof(__half2) : sizeof(float2)) * (size_t)c->nspec;
  for (int off = 0; off < nb; off += c->fx_nb) {
    const int n = std::min(c->fx_nb, nb - off);
    void* Tc = reinterpret_cast<char*>(T) + plane_bytes * off;
    PRC_CHECK(c, timed(c, "fx_filter", [&] {
                if (matrix_x(c) && use_split(c->p))
                  return launch_fx_split(N, c->d_S, c->d_D0, c->d_DN, static_cast<float2*>(Tc), host_dk(c), invN3,
                                         c->d_rc, r0 + off, n, c->d_tw_full, c->wt, c->d_mats, s);
                if (matrix_x(c))
                  return launch_fx_mfma(N, c->d_S, c->d_D0, c->d_DN, Tc, host_dk(c), invN3, c->d_rc, r0 + off, n,
                                        c->d_st_full, c->wt, c->d_mats, s);
                return launch_fx(N, c->d_S, c->d_D0, c->d_DN, Tc, half, host_dk(c), invN3, c->d_rc, r0 + off, n,
                                 c->d_st_fx, c->wt, s);
              }),
              "fused x pass");
  }
  return PRC_D2Z_OK;
}

// The y pass and the z pass (c2r + barrier + max) of one batch over T, on s
int fused_yz(prc_d2z_ctx* c, int r0, int nb, float2* T, hipStream_t s) {
  const int N = c->N;
  const bool half = half_storage(c->p);
  const MixedScales ms = half ? mixed_scales(N) : MixedScales{1.0f, 1.0f, 1.0f};
  TileMap ym = tile_map(N, Axis::kYInPlace);
  ym.nColBlocks *= nb;  // the nb intermediates are consecutive: one launch covers all
  PRC_CHECK(c, timed(c, "inv_y", [&] {
              return half ? launch_tile_h(N, T, ym, ms.s2, c->d_st_full, s)
                          : launch_tile(N, +1, T, T, ym, c->d_st_full, s);
            }),
            "inverse y pass");
  PRC_CHECK(c, timed(c, "fz_barrier", [&] {
              float* rb = c->p.RecordBubbleSizes ? c->d_rb : nullptr;
              if (matrix_z(c) && use_split(c->p))
                return launch_fz_split(N, T, c->d_zre, rb, c->d_rc, r0, nb, c->d_part, c->nblk_fz, c->zscale,
                                       c->d_tw_full, c->d_mats, s);
              if (matrix_z(c))
                return launch_fz_mfma(N, T, c->d_zre, rb, c->d_rc, r0, nb, c->d_part, c->nblk_fz, ms.inv,
                                      c->d_tw_full, c->d_mats, s);
              return launch_fz(N, T, half, c->d_zre, rb, c->d_rc, r0, nb, c->d_part, c->nblk_fz, ms.inv,
                               c->d_st_half, c->d_tw_full, fp32_barrier(c->p), s);
            }),
            "fused z pass");
  return PRC_D2Z_OK;
}

// All fused batches, in order on one stream
int fused_serial(prc_d2z_ctx* c) {
  const int NS = c->p.Nscales;
  int rc = PRC_D2Z_OK;
  for (int r0 = 0; r0 < NS; r0 += c->nbatch) {
    const int nb = std::min(c->nbatch, NS - r0);
    if ((rc = fused_x(c, r0, nb, c->d_T, c->stream)) || (rc = fused_yz(c, r0, nb, c->d_T, c->stream))) return rc;
  }
  return rc;
}

// All launches of one run (forward transform through the last radius).
int enqueue_run(prc_d2z_ctx* c) {
  hipStream_t s = c->stream;
  int rc = forward(c, c->sc.DInit);
  if (rc) return rc;
  PRC_CHECK(c, launch_fill(c->d_zre, c->ncell, 0.0f, s), "zreion init");
  if (c->p.RecordBubbleSizes) PRC_CHECK(c, launch_fill(c->d_rb, c->ncell, 0.0f, s), "rbubble init");
  const int NS = c->p.Nscales;
  const double invN3 = 1.0 / ((double)c->N * c->N * c->N);
  if (c->p.mode == PRC_D2Z_PLAIN_FP32) {
    for (int r = 0; r < NS; ++r)
      if ((rc = plain_radius(c, r, true))) return rc;
    PRC_CHECK(c, launch_moments_finish(c->d_mom, kMomentBlocks, NS, invN3, c->d_sigma_grid, 0, s), "sigma_grid");
  } else {
    if ((rc = fused_serial(c))) return rc;
    PRC_CHECK(c, launch_sigma_finish(c->d_part, c->nblk_fz, NS, invN3, c->d_sigma_grid, s), "sigma_grid");
  }
  return PRC_D2Z_OK;
}

// One run.  With timing on, launches are enqueued directly with events around
// each; otherwise the whole sequence is captured once into a hipGraph (all
// buffers and launch arguments are fixed for the context's lifetime) and
// replayed.
int run_all(prc_d2z_ctx* c) {
  if (!c->have_delta) return c->fail(PRC_D2Z_ERR_STATE, "no density field: call prc_d2z_set_delta first");
  hipStream_t s = c->stream;
  int rc = PRC_D2Z_OK;
  if (c->timing) {
    c->timer.reset();
    if (!c->timer.t0) {
      (void)hipEventCreate(&c->timer.t0);
      (void)hipEventCreate(&c->timer.t1);
    }
    (void)hipEventRecord(c->timer.t0, s);
    rc = enqueue_run(c);
    (void)hipEventRecord(c->timer.t1, s);
  } else if (c->use_graph) {
    if (!c->graph_exec) {
      PRC_CHECK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "graph capture");
      rc = enqueue_run(c);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(s, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return c->hip_fail(e, "graph capture");
      e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        c->graph_exec = nullptr;
        return c->hip_fail(e, "graph instantiate");
      }
    }
    PRC_CHECK(c, hipGraphLaunch(c->graph_exec, s), "graph launch");
  } else {
    rc = enqueue_run(c);
  }
  if (rc) return rc;
  c->have_spectrum = true;
  c->have_result = true;
  return PRC_D2Z_OK;
}

int copy_field_out(prc_d2z_ctx* c, const float* dsrc, float* out, int layout) {
  const int N = c->N;
  if (layout == PRC_D2Z_DENSE) {
    PRC_CHECK(c, hipMemcpyAsync(out, dsrc, sizeof(float) * c->ncell, hipMemcpyDeviceToHost, c->stream), "copy out");
  } else if (layout == PRC_D2Z_PADDED) {
    PRC_CHECK(c, hipMemcpy2DAsync(out, sizeof(float) * (N + 2), dsrc, sizeof(float) * N, sizeof(float) * N,
                                  (size_t)N * N, hipMemcpyDeviceToHost, c->stream),
              "copy out");
  } else {
    return c->fail(PRC_D2Z_ERR_BAD_ARG, "unknown layout");
  }
  PRC_CHECK(c, hipStreamSynchronize(c->stream), "copy out");
  if (layout == PRC_D2Z_PADDED)
    for (long long row = 0; row < (long long)N * N; ++row) out[row * (N + 2) + N] = out[row * (N + 2) + N + 1] = 0.f;
  return PRC_D2Z_OK;
}

}  // namespace

// ================================================= multi-box groups
// Boxes of several sizes run together: one persistent launch per pass covers
// every box (the work lists of the MBox kernels), and the whole step is one
// captured graph on the group's stream.  The boxes' own contexts hold the
// buffers; results land in them as after prc_d2z_run.
struct prc_d2z_group {
  std::vector<prc_d2z_ctx*> ctx;
  int device = 0;
  hipStream_t stream = nullptr;
  MBox* d_boxes = nullptr;  // sorted by N: each size class is a contiguous range
  int NS = 0, nbatch = 1;
  bool rb = false;
  struct Launch {
    MBPass pass;
    const long long* first;
    long long nitems;
    int r0, nb, toff, nblk;
    int b0, nbox, N, maxM;  // the boxes [b0, b0 + nbox) of the launch; N, maxM: x and z passes
    int cls;                // chain: the stream of the launch
  };
  std::vector<Launch> seq;
  std::vector<hipStream_t> cstreams;  // one per chain, forked from `stream` in the step graph
  std::vector<hipEvent_t> events;     // the fork, then one join per class
  std::vector<void*> allocs;
  hipGraphExec_t graph = nullptr;
  std::string err;
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

namespace {
thread_local std::string g_group_error;

// grid of a group launch other than the z pass: one block per work item (the
// blocks of the chains' concurrent launches interleave on the CUs)
constexpr long long kGroupMaxBlocks = 1 << 20;
// chains (streams) of a group step: the hardware queues of a process
constexpr int kGroupChains = 4;

// item offsets of pass (with nb radii) over the boxes [b0, b1), on the device
hipError_t group_first(prc_d2z_group* g, MBPass pass, int nb, int b0, int b1, const long long** out,
                       long long* nitems) {
  std::vector<long long> f(b1 - b0 + 1, 0);
  for (int b = b0; b < b1; ++b) f[b - b0 + 1] = f[b - b0] + mbox_items(pass, g->ctx[b]->N, nb);
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, sizeof(long long) * f.size());
  if (e != hipSuccess) return e;
  g->allocs.push_back(d);
  e = hipMemcpy(d, f.data(), sizeof(long long) * f.size(), hipMemcpyHostToDevice);
  *out = static_cast<const long long*>(d);
  *nitems = f.back();
  return e;
}
}  // namespace

// ======================================== GetHistory without allocation
// Device scratch of prc_d2z_history_noalloc: the integer counts of the
// NHISTORY bins and the lowest / highest occupied bin (the same buffers the
// context-based prc_d2z_history keeps in its context).  Each GPU holds its own
// copy for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returns for that device (the array's name is a
// host-side stand-in, not a device address).
constexpr int kNoallocWords = kNHistory + 64;  // counts, then the bin range (256 B aligned)
static __device__ __align__(256) unsigned int scratch[kNoallocWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on device dev (current), looked up once per device
hipError_t scratch_address(int dev, unsigned int** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<unsigned int*>(g_scratch_ptr[dev]);
  return hipSuccess;
}
}  // namespace

// ===================================================================== ABI
extern "C" {

int prc_d2z_abi_version(void) { return PRC_D2Z_ABI_VERSION; }

void prc_d2z_default_params(prc_d2z_params* p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->Omegam = 0.31f;
  p->Omegab = 0.049f;
  p->Omegal = 0.69f;
  p->h = 0.68f;
  p->ns = 0.965f;
  p->sigma8 = 0.81f;
  p->w = -1.0f;
  p->BoxSize = 200.0f;
  p->zInit = 0.0f;
  p->Rmin = 1.0f;
  p->N = 128;
  p->Nscales = 32;
  p->RecordBubbleSizes = 0;
  p->Rmax = 256.0f;
  p->zeta = 10.0f;
  p->M_min = 1e8f;
  p->mode = PRC_D2Z_FUSED_FP32;
  p->radii_per_launch = 0;
  p->device = 0;
}

const char* prc_d2z_last_error(const prc_d2z_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

size_t prc_d2z_device_bytes(const prc_d2z_ctx* ctx) { return ctx ? ctx->bytes : 0; }

void prc_d2z_destroy(prc_d2z_ctx* c) {
  if (!c) return;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(c->p.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->timer.destroy();
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  for (void* q : c->allocs) (void)hipFree(q);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  (void)hipSetDevice(dev);
  delete c;
}

int prc_d2z_create(const prc_d2z_params* p, prc_d2z_ctx** out) {
  g_create_error.clear();
  if (!p || !out) {
    g_create_error = "null argument";
    return PRC_D2Z_ERR_BAD_ARG;
  }
  *out = nullptr;
  if (!supported_n(p->N)) {
    g_create_error = "unsupported N=" + std::to_string(p->N) +
                     " (supported: 16, 32, 64, 128, 192, 256, 512, 1024, 2048)";
    return PRC_D2Z_ERR_UNSUPPORTED_N;
  }
  if (p->Nscales < 2) {
    g_create_error = "Nscales must be >= 2 (the radius spacing divides by Nscales-1)";
    return PRC_D2Z_ERR_BAD_ARG;
  }
  if (!(p->h > 0) || !(p->BoxSize > 0) || !(p->Rmin > 0) || !(p->Rmax > 0) || !(p->zeta > 1) || !(p->M_min > 0) ||
      !(p->Omegam > 0) || !(p->sigma8 > 0)) {
    g_create_error = "parameters out of range (need h, BoxSize, Rmin, Rmax, M_min, Omegam, sigma8 > 0 and zeta > 1)";
    return PRC_D2Z_ERR_BAD_ARG;
  }
  if (p->mode != PRC_D2Z_PLAIN_FP32 && p->mode != PRC_D2Z_FUSED_FP32 && p->mode != PRC_D2Z_FUSED_MIXED &&
      p->mode != PRC_D2Z_FUSED_HALF) {
    g_create_error = "unknown mode";
    return PRC_D2Z_ERR_BAD_ARG;
  }
  if (p->radii_per_launch < 0 || p->radii_per_launch > kFusedMaxRadii) {
    g_create_error = "radii_per_launch must be in 0.." + std::to_string(kFusedMaxRadii);
    return PRC_D2Z_ERR_BAD_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    g_create_error = "no HIP device";
    return PRC_D2Z_ERR_NO_DEVICE;
  }
  if (p->device < 0 || p->device >= ndev) {
    g_create_error = "device ordinal out of range";
    return PRC_D2Z_ERR_BAD_ARG;
  }
  prc_d2z_ctx* c = new (std::nothrow) prc_d2z_ctx();
  if (!c) {
    g_create_error = "host allocation failed";
    return PRC_D2Z_ERR_OOM;
  }
  c->p = *p;
  c->N = p->N;
  c->Nh = p->N / 2;
  c->ncell = (long long)p->N * p->N * p->N;
  c->nspec = (long long)p->N * p->N * c->Nh;
  // radii per fused batch; the plain path uses two buffers (filtered spectrum
  // / real field, and the intermediate), and at least two are always held
  c->nbatch = (p->mode == PRC_D2Z_PLAIN_FP32)
                  ? 1
                  : std::min(p->Nscales, p->radii_per_launch ? p->radii_per_launch : kDefaultRadiiPerLaunch);
  auto bail = [&](int code, const std::string& m) {
    g_create_error = m;
    prc_d2z_destroy(c);
    return code;
  };
  if (hipSetDevice(p->device) != hipSuccess) return bail(PRC_D2Z_ERR_HIP, "hipSetDevice failed");
  // memory budget check before allocating
  const int NS = p->Nscales;
  if (has_matrix_x(*p)) c->xplan = PRC_D2Z_PLAN_MATRIX;
  if (has_matrix_z(*p)) c->zplan = PRC_D2Z_PLAN_MATRIX;
  c->nblk_fz = default_fz_blocks(*p, c->zplan);
  // intermediates: the fused batch (FP32 or FP16 storage), and at least two
  // FP32-sized buffers for the forward scratch, the plain path and smoothed_field
  const size_t tstore = half_storage(*p) ? sizeof(__half2) : sizeof(float2);
  const size_t fixed = sizeof(float) * c->ncell * (p->RecordBubbleSizes ? 3 : 2) + sizeof(float2) * c->nspec +
                       sizeof(float2) * 2 * (size_t)p->N * p->N + (64u << 20);
  size_t freeb = 0, totalb = 0;
  const bool have_info = hipMemGetInfo(&freeb, &totalb) == hipSuccess;
  auto tbytes_for = [&](int nb) { return std::max((size_t)2 * sizeof(float2) * c->nspec, tstore * c->nspec * nb); };
  // default batch: the largest radii-per-launch (<= 32) whose intermediates fit
  if (p->mode != PRC_D2Z_PLAIN_FP32 && !p->radii_per_launch && have_info)
    while (c->nbatch > 1 && fixed + tbytes_for(c->nbatch) > freeb) --c->nbatch;
  c->tbytes = tbytes_for(c->nbatch);
  const size_t need = fixed + c->tbytes;
  if (have_info && need > freeb)
    return bail(PRC_D2Z_ERR_OOM, "device memory: need " + std::to_string(need >> 20) + " MiB, free " +
                                     std::to_string(freeb >> 20) + " MiB");
  c->nbatch_cap = (p->mode == PRC_D2Z_PLAIN_FP32) ? 1 : (int)std::min<size_t>(kFusedMaxRadii, c->tbytes / (tstore * c->nspec));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(PRC_D2Z_ERR_HIP, "stream creation failed");
  hipError_t e = hipSuccess;
#define A_(ptr, n) \
  if (e == hipSuccess) e = c->alloc(&(ptr), (n))
  A_(c->d_delta, sizeof(float) * c->ncell);
  A_(c->d_S, sizeof(float2) * c->nspec);
  A_(c->d_D0, sizeof(float2) * (size_t)p->N * p->N);
  A_(c->d_DN, sizeof(float2) * (size_t)p->N * p->N);
  A_(c->d_T, c->tbytes);
  A_(c->d_zre, sizeof(float) * c->ncell);
  if (p->RecordBubbleSizes) A_(c->d_rb, sizeof(float) * c->ncell);
  A_(c->d_tw_full, sizeof(float2) * p->N);
  A_(c->d_st_full, sizeof(float2) * std::max(1, stage_twiddle_size(p->N, false)));
  A_(c->d_st_half, sizeof(float2) * std::max(1, stage_twiddle_size(p->N, true)));
  A_(c->d_st_fx, sizeof(float2) * std::max(1, fx_twiddle_size(p->N)));
  const size_t mat_bytes = use_split(*p) ? split_matrix_bytes(p->N) : mfma_matrix_bytes(p->N, half_storage(*p));
  if (mat_bytes) A_(c->d_mats, mat_bytes);
  A_(c->d_mom, sizeof(double2) * (size_t)NS * kMomentBlocks);
  A_(c->d_part, sizeof(double) * (size_t)NS * kMaxZBlocks);
  A_(c->d_sigma_grid, sizeof(float) * NS);
  A_(c->d_sigma, sizeof(float) * NS);
  A_(c->d_rc, sizeof(RadiusConst) * NS);
  A_(c->d_sc, sizeof(SetupScalars));
  A_(c->d_integ, sizeof(double) * (NS + 2));
  A_(c->d_Rint, sizeof(double) * (NS + 2));
  A_(c->d_R, sizeof(float) * NS);
  A_(c->d_hist, sizeof(unsigned) * kNHistory);
  A_(c->d_hrange, sizeof(int) * 2);
  A_(c->d_pk, sizeof(double) * 4 * kPkMaxBins);
  A_(c->d_flag, sizeof(int));
  A_(c->d_maxabs, sizeof(unsigned));
  A_(c->d_mfp_sum, sizeof(double) * kNXMfp);
  A_(c->d_mfp_cnt, sizeof(unsigned) * kNXMfp * kNMfp);
  A_(c->d_zcur, sizeof(float) * kNXMfp);
  A_(c->d_hin, sizeof(float) * kNHistory);
  A_(c->d_chist, sizeof(float) * kNCHistory);
  A_(c->d_zvz, sizeof(float) * kNHistory);
  A_(c->d_grfA, sizeof(double));
#undef A_
  if (e != hipSuccess) return bail(PRC_D2Z_ERR_OOM, std::string("device allocation failed: ") + hipGetErrorString(e));
  // twiddle tables and the FP64 scalar set-up, all on the device
  SetupParams sp{p->Omegam, p->Omegab, p->Omegal, p->h,    p->ns,   p->sigma8, p->w, p->BoxSize,
                 p->zInit,  p->Rmin,   p->Rmax,   p->zeta, p->M_min, p->N,      p->Nscales};
  if ((e = launch_twiddles(c->d_tw_full, p->N, c->stream)) != hipSuccess ||
      (e = launch_stage_twiddles(p->N, c->d_st_full, c->d_st_half, c->stream)) != hipSuccess ||
      (e = launch_fx_twiddles(p->N, c->d_st_fx, c->stream)) != hipSuccess ||
      (mat_bytes && (e = use_split(*p) ? launch_split_matrices(c->d_mats, c->stream)
                                        : launch_mfma_matrices(c->d_mats, c->stream)) != hipSuccess) ||
      (e = launch_setup(sp, c->d_integ, c->d_R, c->d_Rint, c->d_sc, c->d_rc, c->d_sigma, c->stream)) != hipSuccess)
    return bail(PRC_D2Z_ERR_HIP, std::string("set-up kernels: ") + hipGetErrorString(e));
  c->rc.resize(NS);
  if ((e = hipMemcpyAsync(&c->sc, c->d_sc, sizeof(SetupScalars), hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(c->rc.data(), c->d_rc, sizeof(RadiusConst) * NS, hipMemcpyDeviceToHost, c->stream)) !=
          hipSuccess ||
      (e = hipStreamSynchronize(c->stream)) != hipSuccess)
    return bail(PRC_D2Z_ERR_HIP, std::string("set-up copy: ") + hipGetErrorString(e));
  // W(x) table for the fused x pass, covering x = |k| R up to the largest
  // radius and the corner of the k grid (with a small margin)
  {
    float rmax = 0.0f;
    for (const RadiusConst& r : c->rc) rmax = std::max(rmax, r.R);
    const double xmax = (double)host_dk(c) * rmax * std::sqrt(3.0) * (p->N / 2) * 1.001 + 1.0;
    double h = 0.0;
    const int M = wtable_intervals(xmax, use_split(*p) ? kWTableSplitMax : kWTableMax, &h);
    if (M > 0) {
      if ((e = c->alloc(&c->d_wcoef, sizeof(float4) * M)) != hipSuccess)
        return bail(PRC_D2Z_ERR_OOM, "device allocation failed: W table");
      if ((e = launch_wtable(c->d_wcoef, M, h, c->stream)) != hipSuccess ||
          (e = hipStreamSynchronize(c->stream)) != hipSuccess)
        return bail(PRC_D2Z_ERR_HIP, std::string("W table: ") + hipGetErrorString(e));
      c->wt = WTable{c->d_wcoef, M, (float)(1.0 / h)};
    }
  }
  *out = c;
  return PRC_D2Z_OK;
}

// The split z pass's input scale from the new field: a power of two <= 256
// DInit / max|delta| (its stage-A operands stay below 2^16 in FP16; see
// mfma_split_kernels.hip), at most 2^8.  A changed scale drops the captured
// graph, whose launch arguments hold the old one.
static int update_zscale(prc_d2z_ctx* c) {
  if (!use_split(c->p)) return PRC_D2Z_OK;
  unsigned bits = 0;
  PRC_CHECK(c, launch_maxabs(c->d_delta, c->ncell, c->d_maxabs, c->stream), "input range");
  PRC_CHECK(c, hipMemcpyAsync(&bits, c->d_maxabs, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream), "input range");
  PRC_CHECK(c, hipStreamSynchronize(c->stream), "input range");
  float mx = 0.0f;
  memcpy(&mx, &bits, sizeof(mx));
  int e = 8;
  if (mx > 0.0f) e = std::min(8, std::max(-100, (int)std::floor(std::log2(256.0 * (double)c->sc.DInit / mx))));
  const float zs = ldexpf(1.0f, e);
  if (zs != c->zscale && c->graph_exec) {
    (void)hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
  }
  c->zscale = zs;
  return PRC_D2Z_OK;
}

static int set_delta_impl(prc_d2z_ctx* c, const float* src, int layout, hipMemcpyKind kind) {
  if (!c || !src) return c ? c->fail(PRC_D2Z_ERR_BAD_ARG, "null field") : PRC_D2Z_ERR_BAD_ARG;
  (void)hipSetDevice(c->p.device);
  const int N = c->N;
  if (layout == PRC_D2Z_DENSE) {
    PRC_CHECK(c, hipMemcpyAsync(c->d_delta, src, sizeof(float) * c->ncell, kind, c->stream), "copy in");
  } else if (layout == PRC_D2Z_PADDED) {
    PRC_CHECK(c, hipMemcpy2DAsync(c->d_delta, sizeof(float) * N, src, sizeof(float) * (N + 2), sizeof(float) * N,
                                  (size_t)N * N, kind, c->stream),
              "copy in");
  } else {
    return c->fail(PRC_D2Z_ERR_BAD_ARG, "unknown layout");
  }
  int bad = 0;
  PRC_CHECK(c, launch_count_nonfinite(c->d_delta, c->ncell, c->d_flag, c->stream), "input check");
  PRC_CHECK(c, hipMemcpyAsync(&bad, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream), "input check");
  PRC_CHECK(c, hipStreamSynchronize(c->stream), "copy in");
  c->have_delta = false;
  c->have_spectrum = false;
  c->have_result = false;
  if (bad) return c->fail(PRC_D2Z_ERR_NAN_INPUT, "density field holds NaN or Inf");
  if (int rc = update_zscale(c)) return rc;
  c->have_delta = true;
  return PRC_D2Z_OK;
}

int prc_d2z_set_delta(prc_d2z_ctx* c, const float* delta, int layout) {
  return set_delta_impl(c, delta, layout, hipMemcpyHostToDevice);
}

int prc_d2z_set_delta_device(prc_d2z_ctx* c, const float* d_delta, int layout) {
  return set_delta_impl(c, d_delta, layout, hipMemcpyDeviceToDevice);
}

int prc_d2z_generate_delta(prc_d2z_ctx* c, uint64_t seed, int white_noise_only) {
  if (!c) return PRC_D2Z_ERR_BAD_ARG;
  (void)hipSetDevice(c->p.device);
  const int N = c->N, Nh = c->Nh;
  hipStream_t s = c->stream;
  const bool was_timing = c->timing;
  c->timing = false;
  c->have_delta = c->have_spectrum = c->have_result = false;
  PRC_CHECK(c, launch_white_noise(c->d_delta, c->ncell, (unsigned long long)seed, s), "white noise");
  if (!white_noise_only) {
    const double om = c->p.Omegam, ob = c->p.Omegab, h = c->p.h, box = c->p.BoxSize;
    GrfSpectrum sp;
    sp.gamma = om * h * exp(-ob * (1.0 + sqrt(2.0 * h) / om));
    sp.ns = c->p.ns;
    sp.sigma8 = c->p.sigma8;
    sp.dk = 2.0 * M_PI / box;
    sp.vcell = pow(box / N, 3.0);
    sp.invN3 = 1.0 / ((double)N * N * N);
    int rc = forward(c, 1.0f);
    if (rc) {
      c->timing = was_timing;
      return rc;
    }
    PRC_CHECK(c, launch_grf_norm(sp, c->d_grfA, s), "field normalisation");
    PRC_CHECK(c, launch_colour(N, c->d_S, c->d_D0, c->d_DN, sp, c->d_grfA, s), "colouring");
    float2* T0 = c->d_T;
    PRC_CHECK(c, launch_tile(N, +1, c->d_S, T0, tile_map(N, Axis::kXStoT), c->d_st_full, s), "field x pass");
    PRC_CHECK(c, launch_tile(N, +1, T0, T0, tile_map(N, Axis::kYInPlace), c->d_st_full, s), "field y pass");
    PRC_CHECK(c, launch_inv_z(N, T0, c->d_delta, c->d_st_half, c->d_tw_full, s), "field z pass");
  }
  PRC_CHECK(c, hipStreamSynchronize(s), "field generation");
  c->timing = was_timing;
  c->have_spectrum = false;
  if (int rc = update_zscale(c)) return rc;
  c->have_delta = true;
  return PRC_D2Z_OK;
}

int prc_d2z_get_delta(prc_d2z_ctx* c, float* out, int layout) {
  if (!c || !out) return c ? c->fail(PRC_D2Z_ERR_BAD_ARG, "null output") : PRC_D2Z_ERR_BAD_ARG;
  if (!c->have_delta) return c->fail(PRC_D2Z_ERR_STATE, "no density field");
  return copy_field_out(c, c->d_delta, out, layout);
}

int prc_d2z_run_async(prc_d2z_ctx* c) {
  if (!c) return PRC_D2Z_ERR_BAD_ARG;
  (void)hipSetDevice(c->p.device);
  return run_all(c);
}

int prc_d2z_sync(prc_d2z_ctx* c) {
  if (!c) return PRC_D2Z_ERR_BAD_ARG;
  PRC_CHECK(c, hipStreamSynchronize(c->stream), "synchronize");
  return PRC_D2Z_OK;
}

int prc_d2z_run(prc_d2z_ctx* c) {
  int rc = prc_d2z_run_async(c);
  if (rc) return rc;
  return prc_d2z_sync(c);
}

int prc_d2z_get_zreion(prc_d2z_ctx* c, float* out, int layout) {
  if (!c || !out) return c ? c->fail(PRC_D2Z_ERR_BAD_ARG, "null output") : PRC_D2Z_ERR_BAD_ARG;
  if (!c->have_result) return c->fail(PRC_D2Z_ERR_STATE, "no result: call prc_d2z_run first");
  return copy_field_out(c, c->d_zre, out, layout);
}

int prc_d2z_get_rbubble(prc_d2z_ctx* c, float* out, int layout) {
  if (!c || !out) return c ? c->fail(PRC_D2Z_ERR_BAD_ARG, "null output") : PRC_D2Z_ERR_BAD_ARG;
  if (!c->p.RecordBubbleSizes) return c->fail(PRC_D2Z_ERR_STATE, "RecordBubbleSizes is 0");
  if (!c->have_result) return c->fail(PRC_D2Z_ERR_STATE, "no result: call prc_d2z_run first");
  return copy_field_out(c, c->d_rb, out, layout);
}

int prc_d2z_zreion_device(prc_d2z_ctx* c, const float** d) {
  if (!c || !d) return PRC_D2Z_ERR_BAD_ARG;
  *d = c->d_zre;
  return PRC_D2Z_OK;
}

int prc_d2z_get_scalars(prc_d2z_ctx* c, prc_d2z_scalars* o) {
  if (!c || !o) return PRC_D2Z_ERR_BAD_ARG;
  o->DInit = c->sc.DInit;
  o->DHigh = c->sc.DHigh;
  o->A = c->sc.A;
  o->pnorm = c->sc.pnorm;
  o->erfizeta = c->sc.erfizeta;
  o->sigma_min = c->sc.sigma_min;
  return PRC_D2Z_OK;
}

int prc_d2z_get_radius_table(prc_d2z_ctx* c, prc_d2z_radius_row* rows, int n) {
  if (!c || !rows || n < c->p.Nscales) return c ? c->fail(PRC_D2Z_ERR_BAD_ARG, "table too small") : PRC_D2Z_ERR_BAD_ARG;
  std::vector<float> sg(c->p.Nscales, 0.f);
  if (c->have_result) {
    PRC_CHECK(c, hipMemcpyAsync(sg.data(), c->d_sigma_grid, sizeof(float) * sg.size(), hipMemcpyDeviceToHost,
                                c->stream),
              "table copy");
    PRC_CHECK(c, hipStreamSynchronize(c->stream), "table copy");
  }
  for (int r = 0; r < c->p.Nscales; ++r) {
    rows[r].R_mpch = c->rc[r].R * c->p.h;
    rows[r].sigma = c->rc[r].sigma;
    rows[r].sigma_grid = sg[r];
  }
  return PRC_D2Z_OK;
}

int prc_d2z_get_radius_diag(prc_d2z_ctx* c, prc_d2z_radius_diag* rows, int n) {
  if (!c || !rows || n < c->p.Nscales) return c ? c->fail(PRC_D2Z_ERR_BAD_ARG, "table too small") : PRC_D2Z_ERR_BAD_ARG;
  std::vector<prc_d2z_radius_row> base(c->p.Nscales);
  const int rc = prc_d2z_get_radius_table(c, base.data(), c->p.Nscales);
  if (rc) return rc;
  for (int r = 0; r < c->p.Nscales; ++r) {
    rows[r].R_mpch = base[r].R_mpch;
    rows[r].sigma = base[r].sigma;
    rows[r].sigma_grid = base[r].sigma_grid;
    rows[r].A = c->rc[r].A;
    // the kernels hold -inf where the reference's square root is NaN
    rows[r].B = std::isinf(c->rc[r].B) ? NAN : c->rc[r].B;
  }
  return PRC_D2Z_OK;
}

// The scalar set-up pinned by the host: sigma_min and sigma(R) (the two GSL
// QAGS results of delta2zreion_d2z.C:24-26,97-100) and erfizeta (the GSL Brent
// result of :21); B_R is rebuilt from them as the set-up kernel builds it.
int prc_d2z_set_scalar_overrides(prc_d2z_ctx* c, const float* sigma, const float* erfizeta) {
  if (!c) return PRC_D2Z_ERR_BAD_ARG;
  const int NS = c->p.Nscales;
  if (sigma)
    for (int r = 0; r <= NS; ++r)
      if (!(sigma[r] >= 0.0f) || std::isinf(sigma[r]))
        return c->fail(PRC_D2Z_ERR_BAD_ARG, "sigma override must be finite and >= 0 (entry " + std::to_string(r) + ")");
  if (erfizeta && !std::isfinite(*erfizeta)) return c->fail(PRC_D2Z_ERR_BAD_ARG, "erfizeta override must be finite");
  (void)hipSetDevice(c->p.device);
  if (sigma) c->sc.sigma_min = sigma[0];
  if (erfizeta) c->sc.erfizeta = *erfizeta;
  std::vector<float> sig(NS);
  for (int r = 0; r < NS; ++r) {
    if (sigma) c->rc[r].sigma = sigma[r + 1];
    const float smin = c->sc.sigma_min, s = c->rc[r].sigma;
    const float dd = smin * smin - s * s;
    c->rc[r].B = barrier_offset(dd, c->sc.erfizeta);
    sig[r] = c->rc[r].sigma;
  }
  PRC_CHECK(c, hipMemcpyAsync(c->d_rc, c->rc.data(), sizeof(RadiusConst) * NS, hipMemcpyHostToDevice, c->stream),
            "overrides");
  PRC_CHECK(c, hipMemcpyAsync(c->d_sigma, sig.data(), sizeof(float) * NS, hipMemcpyHostToDevice, c->stream),
            "overrides");
  PRC_CHECK(c, hipMemcpyAsync(c->d_sc, &c->sc, sizeof(SetupScalars), hipMemcpyHostToDevice, c->stream), "overrides");
  PRC_CHECK(c, hipStreamSynchronize(c->stream), "overrides");
  c->have_result = false;
  return PRC_D2Z_OK;
}

int prc_d2z_history(prc_d2z_ctx* c, uint32_t* counts) {
  if (!c || !counts) return PRC_D2Z_ERR_BAD_ARG;
  if (!c->have_result) return c->fail(PRC_D2Z_ERR_STATE, "no result: call prc_d2z_run first");
  (void)hipSetDevice(c->p.device);
  // dzhistory = (HISTORY_FINAL - HISTORY_INITIAL)/(NHISTORY-1) stored as float
  const float dzf = (float)((0.01 - 49.9) / (kNHistory - 1));
  PRC_CHECK(c, hipMemsetAsync(c->d_hist, 0, sizeof(unsigned) * kNHistory, c->stream), "history");
  PRC_CHECK(c, launch_history(c->d_zre, c->ncell, c->d_hist, (double)dzf, c->d_hrange, c->stream), "history");
  PRC_CHECK(c, hipMemcpyAsync(counts, c->d_hist, sizeof(unsigned) * kNHistory, hipMemcpyDeviceToHost, c->stream),
            "history");
  PRC_CHECK(c, hipStreamSynchronize(c->stream), "history");
  return PRC_D2Z_OK;
}

int prc_d2z_history_noalloc(const float* d_zreion, int64_t ncells, uint32_t* counts) {
  g_create_error.clear();
  if (!d_zreion || !counts) {
    g_create_error = "null argument";
    return PRC_D2Z_ERR_BAD_ARG;
  }
  if (ncells < 1 || ncells > (int64_t)UINT32_MAX) {
    g_create_error = "ncells must be in 1..2^32-1 (32-bit bin counts); got " + std::to_string((long long)ncells);
    return PRC_D2Z_ERR_BAD_ARG;
  }
  if (reinterpret_cast<uintptr_t>(d_zreion) % 16 != 0) {
    g_create_error = "d_zreion must be 16-byte aligned (the kernels read float4)";
    return PRC_D2Z_ERR_BAD_ARG;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    g_create_error = "no current HIP device";
    return PRC_D2Z_ERR_NO_DEVICE;
  }
  // the field must be device memory of the current device: the kernels read it
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, d_zreion) != hipSuccess ||
      (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged) || attr.device != dev) {
    (void)hipGetLastError();
    g_create_error = "d_zreion is not device memory of the current device";
    return PRC_D2Z_ERR_BAD_ARG;
  }
  unsigned int* sp = nullptr;
  if ((e = scratch_address(dev, &sp)) != hipSuccess) {
    g_create_error = std::string("scratch address: ") + hipGetErrorString(e);
    return PRC_D2Z_ERR_HIP;
  }
  // the previous call's counts stay in the array: clear them first
  const float dzf = (float)((0.01 - 49.9) / (kNHistory - 1));
  hipStream_t s = nullptr;
  if ((e = launch_fill(reinterpret_cast<float*>(sp), kNHistory, 0.0f, s)) != hipSuccess ||
      (e = launch_history(d_zreion, ncells, sp, (double)dzf, reinterpret_cast<int*>(sp + kNHistory), s)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(counts, sp, sizeof(unsigned int) * kNHistory, hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess) {
    g_create_error = std::string("history: ") + hipGetErrorString(e);
    return PRC_D2Z_ERR_HIP;
  }
  return PRC_D2Z_OK;
}

// ----------------------------------------------------- history correction
namespace {

constexpr double kHistInitial = 49.9, kHistFinal = 0.01;  // HISTORY_INITIAL/FINAL, allvars_d2z.h:2-3
constexpr double kXMfpInitial = 0.01, kXMfpFinal = 0.999;  // X_MFP_INITIAL/FINAL, allvars_d2z.h:10-11

float history_dz() { return (float)((kHistFinal - kHistInitial) / (kNHistory - 1)); }
float chistory_dz() { return (float)((kHistFinal - kHistInitial) / (kNCHistory - 1)); }
float mfp_dx() { return (float)((kXMfpFinal - kXMfpInitial) / (kNXMfp - 1)); }

// Redshift at which the history first crosses x (bubblehistory_d2z.C:113-125):
// strict bracket history[i] < x < history[i+1], linear in between; below the
// first value 49.9, above the last 0.01; no bracket keeps the previous value
// For a non-decreasing history (a float running sum of counts always is) the
// only bracket is at the last i with history[i] < x, found by bisection.
float history_redshift(const float* history, float x, float prev, float dzh, bool sorted) {
  if (x < history[0]) return (float)kHistInitial;
  if (x > history[kNHistory - 1]) return (float)kHistFinal;
  int i = -1;
  if (sorted) {
    const int q = (int)(std::lower_bound(history, history + kNHistory, x) - history) - 1;
    if (q >= 0 && q < kNHistory - 1 && history[q + 1] > x) i = q;
  } else {
    for (int q = 0; q < kNHistory - 1; ++q)
      if (x > history[q] && x < history[q + 1]) {
        i = q;
        break;
      }
  }
  if (i < 0) return prev;
  const float f = (x - history[i]) / (history[i + 1] - history[i]);
  const double z0 = kHistInitial + (double)((float)i * dzh);
  const double z1 = kHistInitial + (double)((float)(i + 1) * dzh);
  return (float)((double)(1.0f - f) * z0 + (double)f * z1);
}

// Float2FloatTable::Float2Float (tables.C:27-38): table[bin] at min + bin delta,
// delta = (max - min)/n; the last interval holds table[n-1] (the reference
// reads table[n] there)
float table_value(const float* t, int n, float tmin, float tmax, float v) {
  if (v <= tmin) return t[0];
  if (v >= tmax) return t[n - 1];
  const float delta = (tmax - tmin) / (float)n;
  const int bin = (int)((v - tmin) / delta);
  if (bin >= n - 1) return t[n - 1];
  const float f = (v - (tmin + (float)bin * delta)) / delta;
  return (1.0f - f) * t[bin] + f * t[bin + 1];
}

// BubbleMFP (bubblehistory_d2z.C:246-266): mfp at ionized fraction x, linear
float bubble_mfp_at(const float* mfp, float x, float dx) {
  const int itab = (int)(((double)x - kXMfpInitial) / (double)dx);
  if (itab < 0) return mfp[0];
  if (itab >= kNXMfp - 1) return mfp[kNXMfp - 1];
  const float f = (float)(((double)(x - (float)itab * dx) - kXMfpInitial) / (double)dx);
  return (1.0f - f) * mfp[itab] + f * mfp[itab + 1];
}

}  // namespace

int prc_d2z_bubble_mfp(prc_d2z_ctx* c, const float* history, int nskewers, uint64_t seed, float* mfp,
                       float* mfp_x, float* dfdmfp) {
  if (!c || !history || !mfp) return PRC_D2Z_ERR_BAD_ARG;
  if (!c->have_result) return c->fail(PRC_D2Z_ERR_STATE, "no result: call prc_d2z_run first");
  if (nskewers == 0) nskewers = kNSkewers;
  if (nskewers < 0 || nskewers > (1 << 24)) return c->fail(PRC_D2Z_ERR_BAD_ARG, "nskewers must be 0..2^24");
  if (c->N < 3) return c->fail(PRC_D2Z_ERR_UNSUPPORTED_N, "skewers need N >= 3");
  (void)hipSetDevice(c->p.device);
  // zcur per ionized fraction x = 0.01 + ix dxmfp (bubblehistory_d2z.C:104-125)
  const float dzh = history_dz(), dx = mfp_dx();
  std::vector<float> zcur(kNXMfp);
  float prev = (float)kHistInitial;
  const bool sorted = std::is_sorted(history, history + kNHistory);
  for (int ix = 0; ix < kNXMfp; ++ix) {
    const float x = (float)(kXMfpInitial + (double)((float)ix * dx));
    prev = zcur[ix] = history_redshift(history, x, prev, dzh, sorted);
  }
  // CellSize = BoxSize/N; bins of ln distance between ln CellSize and ln BoxSize
  // (bubblehistory_d2z.C:95-101), with BoxSize in Mpc: the d2z program divides
  // the parameter file's Mpc/h by h before anything else (main_d2z.C:29)
  const float box = box_mpc(c->p);
  const float cell = box / (float)c->N;
  SkewerParams sp;
  sp.N = c->N;
  sp.nskewers = nskewers;
  sp.seed_lo = (unsigned)(seed & 0xffffffffu);
  sp.seed_hi = (unsigned)(seed >> 32);
  sp.cell = (double)cell;
  sp.lmin = logf(cell);
  sp.lmax = (float)log(1.0 * (double)box);
  sp.dmfp = (sp.lmax - sp.lmin) / (float)kNMfp;
  hipStream_t s = c->stream;
  PRC_CHECK(c, hipMemcpyAsync(c->d_zcur, zcur.data(), sizeof(float) * kNXMfp, hipMemcpyHostToDevice, s), "skewers");
  PRC_CHECK(c, hipMemsetAsync(c->d_mfp_sum, 0, sizeof(double) * kNXMfp, s), "skewers");
  PRC_CHECK(c, hipMemsetAsync(c->d_mfp_cnt, 0, sizeof(unsigned) * kNXMfp * kNMfp, s), "skewers");
  PRC_CHECK(c, hipMemsetAsync(c->d_flag, 0, sizeof(int), s), "skewers");
  PRC_CHECK(c, launch_skewers(c->d_zre, c->d_zcur, sp, c->d_mfp_sum, c->d_mfp_cnt, c->d_flag, s), "skewers");
  std::vector<double> sums(kNXMfp);
  std::vector<unsigned> cnt((size_t)kNXMfp * kNMfp);
  int failed = 0;
  PRC_CHECK(c, hipMemcpyAsync(sums.data(), c->d_mfp_sum, sizeof(double) * kNXMfp, hipMemcpyDeviceToHost, s),
            "skewers");
  PRC_CHECK(c, hipMemcpyAsync(cnt.data(), c->d_mfp_cnt, sizeof(unsigned) * cnt.size(), hipMemcpyDeviceToHost, s),
            "skewers");
  PRC_CHECK(c, hipMemcpyAsync(&failed, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, s), "skewers");
  PRC_CHECK(c, hipStreamSynchronize(s), "skewers");
  if (failed)
    return c->fail(PRC_D2Z_ERR_BAD_ARG, "skewers: no start cell with zreion > zcur in " +
                                            std::to_string(kMaxStartDraws) + " draws (history does not match the field)");
  // mean over 2 nskewers walks (bubblehistory_d2z.C:236-240)
  for (int ix = 0; ix < kNXMfp; ++ix) {
    mfp[ix] = (float)(sums[ix] / (2.0 * nskewers));
    if (mfp_x) mfp_x[ix] = (float)(kXMfpInitial + (double)((float)ix * dx));
  }
  if (dfdmfp)
    for (size_t q = 0; q < cnt.size(); ++q) dfdmfp[q] = (float)cnt[q];
  return PRC_D2Z_OK;
}

int prc_d2z_corrected_history(prc_d2z_ctx* c, const float* fcoll, int nfcoll, float fmin, float fmax,
                              const float* mfp, float* chistory) {
  if (!c || !fcoll || !mfp || !chistory) return PRC_D2Z_ERR_BAD_ARG;
  if (nfcoll < 2 || !(fmax > fmin)) return c->fail(PRC_D2Z_ERR_BAD_ARG, "fcoll table: need n >= 2 and max > min");
  // GetCHistory (bubblehistory_d2z.C:44-80), float arithmetic as the reference
  const float dzc = chistory_dz(), dx = mfp_dx();
  const float rmfp = c->p.Rmax, zeta = c->p.zeta;
  chistory[0] = 0.0f;
  for (int i = 1; i < kNCHistory; ++i) {
    const float x1 = chistory[i - 1];
    const float z1 = (float)(kHistInitial + (double)((float)i * dzc));
    const float z2 = (float)(kHistInitial + (double)((float)(i + 1) * dzc));
    const float f1 = table_value(fcoll, nfcoll, fmin, fmax, z1);
    const float f2 = table_value(fcoll, nfcoll, fmin, fmax, z2);
    const float df = f2 - f1;
    const float rb = bubble_mfp_at(mfp, x1, dx);
    float v = chistory[i - 1] + df / (1.0f + rb / rmfp) * zeta;
    if (v >= 1.0f) v = 1.0f;
    chistory[i] = v;
  }
  return PRC_D2Z_OK;
}

int prc_d2z_correct_zreion(prc_d2z_ctx* c, const float* history, const float* chistory) {
  if (!c || !history || !chistory) return PRC_D2Z_ERR_BAD_ARG;
  if (!c->have_result) return c->fail(PRC_D2Z_ERR_STATE, "no result: call prc_d2z_run first");
  (void)hipSetDevice(c->p.device);
  hipStream_t s = c->stream;
  std::vector<float> zvz(kNHistory);
  PRC_CHECK(c, hipMemcpyAsync(c->d_hin, history, sizeof(float) * kNHistory, hipMemcpyHostToDevice, s), "correct");
  PRC_CHECK(c, hipMemcpyAsync(c->d_chist, chistory, sizeof(float) * kNCHistory, hipMemcpyHostToDevice, s),
            "correct");
  PRC_CHECK(c, launch_zvz(c->d_hin, c->d_chist, chistory_dz(), c->d_zvz, s), "correct");
  PRC_CHECK(c, hipMemcpyAsync(zvz.data(), c->d_zvz, sizeof(float) * kNHistory, hipMemcpyDeviceToHost, s), "correct");
  PRC_CHECK(c, hipStreamSynchronize(s), "correct");
  // bins without a value take the next bin's (CorrectHistory, :316)
  for (int i = 0; i < kNHistory; ++i)
    if (std::isnan(zvz[i]))
      return c->fail(PRC_D2Z_ERR_BAD_ARG, "corrected history brackets no value of the history (bin " +
                                              std::to_string(i) + ")");
  for (int i = kNHistory - 2; i >= 0; --i)
    if (zvz[i] < 0.0f) zvz[i] = zvz[i + 1];
  PRC_CHECK(c, hipMemcpyAsync(c->d_zvz, zvz.data(), sizeof(float) * kNHistory, hipMemcpyHostToDevice, s), "correct");
  PRC_CHECK(c, launch_remap(c->d_zre, c->ncell, c->d_zvz, (double)history_dz(), s), "correct");
  PRC_CHECK(c, hipStreamSynchronize(s), "correct");
  return PRC_D2Z_OK;
}

int prc_d2z_smoothed_field(prc_d2z_ctx* c, int r, float* out) {
  if (!c || !out) return PRC_D2Z_ERR_BAD_ARG;
  if (r < 0 || r >= c->p.Nscales) return c->fail(PRC_D2Z_ERR_BAD_ARG, "radius index out of range");
  if (!c->have_delta) return c->fail(PRC_D2Z_ERR_STATE, "no density field: call prc_d2z_set_delta first");
  (void)hipSetDevice(c->p.device);
  const bool was_timing = c->timing;
  c->timing = false;
  int rc = c->have_spectrum ? PRC_D2Z_OK : forward(c, c->sc.DInit);
  if (!rc) {
    if (c->p.mode == PRC_D2Z_PLAIN_FP32) {
      rc = plain_radius(c, r, false);
    } else {
      const int N = c->N;
      const bool half = half_storage(c->p);
      const MixedScales ms = half ? mixed_scales(N) : MixedScales{1.0f, 1.0f, 1.0f};
      const float invN3 = (float)(1.0 / ((double)N * N * N)) * ms.s1;
      // the real output goes behind the first intermediate (FP32 sized region)
      float* real = reinterpret_cast<float*>(c->d_T + c->nspec);
      const TileMap ym = tile_map(N, Axis::kYInPlace);
      // the mixed mode's x pass is its split matrix-core kernel; its z pass
      // fuses the barrier, so delta_R comes from the FP32 z c2r here
      hipError_t e = (matrix_x(c) && use_split(c->p))
                         ? launch_fx_split(N, c->d_S, c->d_D0, c->d_DN, c->d_T, host_dk(c), invN3, c->d_rc, r, 1,
                                           c->d_tw_full, c->wt, c->d_mats, c->stream)
                     : matrix_x(c)
                         ? launch_fx_mfma(N, c->d_S, c->d_D0, c->d_DN, c->d_T, host_dk(c), invN3, c->d_rc, r, 1,
                                          c->d_st_full, c->wt, c->d_mats, c->stream)
                         : launch_fx(N, c->d_S, c->d_D0, c->d_DN, c->d_T, half, host_dk(c), invN3, c->d_rc, r, 1,
                                     c->d_st_fx, c->wt, c->stream);
      if (e == hipSuccess)
        e = half ? launch_tile_h(N, c->d_T, ym, ms.s2, c->d_st_full, c->stream)
                 : launch_tile(N, +1, c->d_T, c->d_T, ym, c->d_st_full, c->stream);
      if (e == hipSuccess)
        e = half ? launch_inv_z_h(N, c->d_T, real, ms.inv, c->d_st_half, c->d_tw_full, c->stream)
                 : launch_inv_z(N, c->d_T, real, c->d_st_half, c->d_tw_full, c->stream);
      if (e != hipSuccess) rc = c->hip_fail(e, "smoothed field");
    }
  }
  c->timing = was_timing;
  if (rc) return rc;
  const float* real = reinterpret_cast<const float*>(c->p.mode == PRC_D2Z_PLAIN_FP32 ? c->d_T : c->d_T + c->nspec);
  return copy_field_out(c, real, out, PRC_D2Z_DENSE);
}

// Delta2Velocity (delta2velocity_map.C:8-97): psi_a = IFFT(i k_a / k^2 FFT(delta))
// with the reference's N^-1.5 after each transform, i.e. 1/N^3 in all, and
// k = n 2 pi / BoxSize (BoxSize as given).  The context's spectrum is that of
// delta / DInit (the d2z input scaling), so DInit is folded into the scale.
int prc_d2z_displacement(prc_d2z_ctx* c, int axis, float* out, int layout) {
  if (!c || !out) return PRC_D2Z_ERR_BAD_ARG;
  if (axis < 0 || axis > 2) return c->fail(PRC_D2Z_ERR_BAD_ARG, "axis must be 0 (x), 1 (y) or 2 (z)");
  if (layout != PRC_D2Z_DENSE && layout != PRC_D2Z_PADDED) return c->fail(PRC_D2Z_ERR_BAD_ARG, "unknown layout");
  if (!c->have_delta) return c->fail(PRC_D2Z_ERR_STATE, "no density field: call prc_d2z_set_delta first");
  (void)hipSetDevice(c->p.device);
  const bool was_timing = c->timing;
  c->timing = false;
  int rc = c->have_spectrum ? PRC_D2Z_OK : forward(c, c->sc.DInit);
  c->timing = was_timing;
  if (rc) return rc;
  const int N = c->N;
  hipStream_t s = c->stream;
  float2* T1 = c->d_T;
  float2* T2 = c->d_T + c->nspec;  // the intermediates hold at least two FP32 spectra
  const TileMap xm = tile_map(N, Axis::kXStoT), ym = tile_map(N, Axis::kYInPlace);
  const float dk = (float)(2.0 * M_PI / (double)c->p.BoxSize);
  const float scale = (float)((double)c->sc.DInit / ((double)N * N * N));
  PRC_CHECK(c, launch_disp(N, c->d_S, c->d_D0, c->d_DN, T1, axis, dk, scale, s), "displacement multiply");
  PRC_CHECK(c, launch_tile(N, +1, T1, T2, xm, c->d_st_full, s), "inverse x pass");
  PRC_CHECK(c, launch_tile(N, +1, T2, T2, ym, c->d_st_full, s), "inverse y pass");
  float* real = reinterpret_cast<float*>(T1);
  PRC_CHECK(c, launch_inv_z(N, T2, real, c->d_st_half, c->d_tw_full, s), "inverse z pass");
  return copy_field_out(c, real, out, layout);
}

// PowerSpectrum (powerspectrum_lmb.C:8-139) of the context's density field:
// the set-up in the reference's float arithmetic (dk, kmax, d3k, dlnk), the
// binning on the GPU, the per-bin normalisations as the reference's
// (:108-110), NaN for empty bins as its divisions by zero give.
int prc_d2z_power_spectrum(prc_d2z_ctx* c, int nk, float kmin, float kmax, double* ps, double* ook2, double* kmean,
                           int64_t* nmodes) {
  if (!c || !ps || !ook2 || !kmean || !nmodes) return PRC_D2Z_ERR_BAD_ARG;
  if (nk < 1 || nk > kPkMaxBins)
    return c->fail(PRC_D2Z_ERR_BAD_ARG, "nk must be in 1.." + std::to_string(kPkMaxBins));
  if (!(kmin > 0.0f) || !(kmax > kmin)) return c->fail(PRC_D2Z_ERR_BAD_ARG, "need 0 < kmin < kmax");
  if (!c->have_delta) return c->fail(PRC_D2Z_ERR_STATE, "no density field: call prc_d2z_set_delta first");
  (void)hipSetDevice(c->p.device);
  const bool was_timing = c->timing;
  c->timing = false;
  int rc = c->have_spectrum ? PRC_D2Z_OK : forward(c, c->sc.DInit);
  c->timing = was_timing;
  if (rc) return rc;
  const int N = c->N;
  const float dk = (float)(2. * M_PI / (double)c->p.BoxSize);  // powerspectrum_lmb.C:18
  const float kmaxg = dk * N;                                    // :19
  const float d3k = (float)pow((double)kmaxg, 3);                // :20
  const float lnkmin = logf(kmin);
  const float dlnk = (logf(kmax) - logf(kmin)) / nk;              // :26
  // |FFT(delta) N^-1.5|^2 / d3k from S = FFT(delta / DInit)
  const double pscale = (double)c->sc.DInit * c->sc.DInit / ((double)N * N * N) / (double)d3k;
  unsigned long long* d_count = reinterpret_cast<unsigned long long*>(c->d_pk);
  double* d_sums = c->d_pk + kPkMaxBins;
  PRC_CHECK(c, launch_pk(N, c->d_S, c->d_D0, c->d_DN, dk, lnkmin, dlnk, nk, pscale, d_count, d_sums, c->stream),
            "power spectrum");
  std::vector<unsigned long long> cnt(nk);
  std::vector<double> sums(3 * (size_t)nk);
  PRC_CHECK(c, hipMemcpyAsync(cnt.data(), d_count, sizeof(unsigned long long) * nk, hipMemcpyDeviceToHost, c->stream),
            "power spectrum copy");
  PRC_CHECK(c, hipMemcpyAsync(sums.data(), d_sums, sizeof(double) * 3 * nk, hipMemcpyDeviceToHost, c->stream),
            "power spectrum copy");
  PRC_CHECK(c, hipStreamSynchronize(c->stream), "power spectrum copy");
  const double twopi3 = pow(2 * M_PI, 3);
  for (int b = 0; b < nk; ++b) {
    const double n = (double)cnt[b];
    nmodes[b] = (int64_t)cnt[b];
    ook2[b] = sums[nk + b] / n;
    kmean[b] = sums[2 * nk + b] / n;
    ps[b] = sums[b] * twopi3 / n;
  }
  return PRC_D2Z_OK;
}

int prc_d2z_set_timing(prc_d2z_ctx* c, int enable) {
  if (!c) return PRC_D2Z_ERR_BAD_ARG;
  c->timing = enable != 0;
  return PRC_D2Z_OK;
}

int prc_d2z_set_launch_config(prc_d2z_ctx* c, int radii_per_launch, int z_blocks) {
  if (!c) return PRC_D2Z_ERR_BAD_ARG;
  if (c->p.mode == PRC_D2Z_PLAIN_FP32) return c->fail(PRC_D2Z_ERR_STATE, "launch configuration applies to fused modes");
  if (radii_per_launch < 1 || radii_per_launch > c->nbatch_cap)
    return c->fail(PRC_D2Z_ERR_BAD_ARG, "radii_per_launch must be in 1.." + std::to_string(c->nbatch_cap));
  if (z_blocks < 1 || z_blocks > kMaxZBlocks)
    return c->fail(PRC_D2Z_ERR_BAD_ARG, "z_blocks must be in 1.." + std::to_string(kMaxZBlocks));
  c->nbatch = std::min(radii_per_launch, c->p.Nscales);
  c->nblk_fz = z_blocks;
  if (c->graph_exec) {
    (void)hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
  }
  return PRC_D2Z_OK;
}

int prc_d2z_get_launch_config(const prc_d2z_ctx* c, int* radii_per_launch, int* z_blocks, int* max_radii) {
  if (!c) return PRC_D2Z_ERR_BAD_ARG;
  if (radii_per_launch) *radii_per_launch = c->nbatch;
  if (z_blocks) *z_blocks = c->nblk_fz;
  if (max_radii) *max_radii = c->nbatch_cap;
  return PRC_D2Z_OK;
}

int prc_d2z_set_plan(prc_d2z_ctx* c, int x_plan, int z_plan, int x_radii_per_launch) {
  if (!c) return PRC_D2Z_ERR_BAD_ARG;
  if (c->p.mode == PRC_D2Z_PLAIN_FP32) return c->fail(PRC_D2Z_ERR_STATE, "plans apply to fused modes");
  for (int pl : {x_plan, z_plan})
    if (pl != PRC_D2Z_PLAN_VECTOR && pl != PRC_D2Z_PLAN_MATRIX)
      return c->fail(PRC_D2Z_ERR_BAD_ARG, "plan must be PRC_D2Z_PLAN_VECTOR or PRC_D2Z_PLAN_MATRIX");
  if ((x_plan == PRC_D2Z_PLAN_MATRIX && !has_matrix_x(c->p)) || (z_plan == PRC_D2Z_PLAN_MATRIX && !has_matrix_z(c->p)))
    return c->fail(PRC_D2Z_ERR_UNSUPPORTED_N,
                   "no matrix-core plan for this pass, mode and N (mixed: N = 1024; half: x at 1024 and 2048, z at 1024)");
  if (x_radii_per_launch < 1 || x_radii_per_launch > kFxMaxRadii)
    return c->fail(PRC_D2Z_ERR_BAD_ARG, "x_radii_per_launch must be in 1.." + std::to_string(kFxMaxRadii));
  if (z_plan != c->zplan) c->nblk_fz = default_fz_blocks(c->p, z_plan);
  c->xplan = x_plan;
  c->zplan = z_plan;
  c->fx_nb = x_radii_per_launch;
  if (c->graph_exec) {
    (void)hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
  }
  return PRC_D2Z_OK;
}

int prc_d2z_get_plan(const prc_d2z_ctx* c, int* x_plan, int* z_plan, int* x_radii_per_launch) {
  if (!c) return PRC_D2Z_ERR_BAD_ARG;
  if (x_plan) *x_plan = c->xplan;
  if (z_plan) *z_plan = c->zplan;
  if (x_radii_per_launch) *x_radii_per_launch = c->fx_nb;
  return PRC_D2Z_OK;
}

// Times the fused batches covering the first min(Nscales, cap) radii for
// each candidate and keeps the fastest per radius.  Candidates: the FFT plans
// of the x and z passes (vector radix plans, and matrix-core radix-16 plans
// where the mode has them), radii per x launch (8, 16), radii per batch (cap
// and cap/2) and the z grid (the plan's resident capacity and twice it).  Every
// candidate computes the same transforms; plans differ only in rounding.
// Needs a density field; leaves zreion undefined until the next run.
int prc_d2z_autotune(prc_d2z_ctx* c, prc_d2z_tuning* out) {
  if (!c || !out) return PRC_D2Z_ERR_BAD_ARG;
  if (c->p.mode == PRC_D2Z_PLAIN_FP32) return c->fail(PRC_D2Z_ERR_STATE, "autotune applies to fused modes");
  if (!c->have_delta) return c->fail(PRC_D2Z_ERR_STATE, "no density field: call prc_d2z_set_delta first");
  (void)hipSetDevice(c->p.device);
  const bool was_timing = c->timing;
  c->timing = false;
  int rc = c->have_spectrum ? PRC_D2Z_OK : forward(c, c->sc.DInit);
  if (rc) {
    c->timing = was_timing;
    return rc;
  }
  const int nr = std::min(c->p.Nscales, c->nbatch_cap);
  std::vector<int> nbs = {c->nbatch_cap};
  if (c->nbatch_cap >= 2) nbs.push_back(c->nbatch_cap / 2);
  std::vector<int> xplans = {PRC_D2Z_PLAN_VECTOR}, zplans = {PRC_D2Z_PLAN_VECTOR};
  if (has_matrix_x(c->p)) xplans.push_back(PRC_D2Z_PLAN_MATRIX);
  if (has_matrix_z(c->p)) zplans.push_back(PRC_D2Z_PLAN_MATRIX);
  const std::vector<int> fxs = {8, kFxMaxRadii};
  hipEvent_t a = nullptr, b = nullptr;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  double best = 1e30;
  prc_d2z_tuning bt{c->nbatch, c->nblk_fz, 0.0, c->xplan, c->zplan, c->fx_nb};
  const prc_d2z_tuning keep = bt;
  for (int xp : xplans)
    for (int zp : zplans)
      for (int fx : fxs)
        for (int nb : nbs) {
          const int g0 = default_fz_blocks(c->p, zp);
          for (int g : {g0, 2 * g0}) {
            if (g > kMaxZBlocks || rc) continue;
            c->xplan = xp;
            c->zplan = zp;
            c->fx_nb = fx;
            c->nbatch = nb;
            c->nblk_fz = g;
            for (int rep = 0; rep < 2 && !rc; ++rep) {  // first repetition warms up
              (void)hipEventRecord(a, c->stream);
              for (int r0 = 0; r0 < nr && !rc; r0 += nb) {
                const int n = std::min(nb, nr - r0);
                if (!(rc = fused_x(c, r0, n, c->d_T, c->stream))) rc = fused_yz(c, r0, n, c->d_T, c->stream);
              }
              (void)hipEventRecord(b, c->stream);
            }
            if (rc) continue;
            (void)hipEventSynchronize(b);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, a, b);
            const double per = ms / nr;
            if (per < best) {
              best = per;
              bt = prc_d2z_tuning{nb, g, per, xp, zp, fx};
            }
          }
        }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  c->timing = was_timing;
  c->have_result = false;
  c->xplan = keep.x_plan;
  c->zplan = keep.z_plan;
  c->fx_nb = keep.x_radii_per_launch;
  c->nbatch = keep.radii_per_launch;
  c->nblk_fz = keep.z_blocks;
  if (rc) return rc;
  *out = bt;
  if ((rc = prc_d2z_set_plan(c, bt.x_plan, bt.z_plan, bt.x_radii_per_launch))) return rc;
  return prc_d2z_set_launch_config(c, bt.radii_per_launch, bt.z_blocks);
}

int prc_d2z_run_many(prc_d2z_ctx* const* ctxs, int n) {
  if (!ctxs || n < 0) return PRC_D2Z_ERR_BAD_ARG;
  for (int k = 0; k < n; ++k) {
    int rc = prc_d2z_run_async(ctxs[k]);
    if (rc) return rc;
  }
  for (int k = 0; k < n; ++k) {
    int rc = prc_d2z_sync(ctxs[k]);
    if (rc) return rc;
  }
  return PRC_D2Z_OK;
}

void prc_d2z_group_destroy(prc_d2z_group* g) {
  if (!g) return;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(g->device);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  if (g->graph) (void)hipGraphExecDestroy(g->graph);
  for (void* q : g->allocs) (void)hipFree(q);
  for (hipEvent_t ev : g->events) (void)hipEventDestroy(ev);
  for (hipStream_t cs : g->cstreams) (void)hipStreamDestroy(cs);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  (void)hipSetDevice(dev);
  delete g;
}

const char* prc_d2z_group_last_error(const prc_d2z_group* g) { return g ? g->err.c_str() : g_group_error.c_str(); }

int prc_d2z_group_create(prc_d2z_ctx* const* ctxs, int n, prc_d2z_group** out) {
  g_group_error.clear();
  if (!ctxs || !out || n < 1 || n > kMBoxMax) {
    g_group_error = "need 1.." + std::to_string(kMBoxMax) + " contexts";
    return PRC_D2Z_ERR_BAD_ARG;
  }
  *out = nullptr;
  const prc_d2z_ctx* c0 = ctxs[0];
  for (int k = 0; k < n; ++k) {
    const prc_d2z_ctx* c = ctxs[k];
    std::string why;
    if (!c) why = "null context";
    else if (c->p.device != c0->p.device) why = "contexts on different devices";
    else if (c->p.mode != PRC_D2Z_FUSED_FP32 && !(c->p.mode == PRC_D2Z_FUSED_MIXED && !use_split(c->p)))
      why = "group boxes run the fused FP32 kernels (mode FUSED_FP32, or FUSED_MIXED below N = 1024)";
    else if (!mbox_supported(c->N)) why = "N = " + std::to_string(c->N) + " (group boxes: N = 64, 128, 192, 256)";
    else if (c->p.Nscales != c0->p.Nscales || c->nbatch != c0->nbatch) why = "Nscales and radii per batch must agree";
    else if (c->p.Nscales > 8 * kFusedMaxRadii) why = "Nscales above 256";
    else if ((c->p.RecordBubbleSizes != 0) != (c0->p.RecordBubbleSizes != 0)) why = "RecordBubbleSizes must agree";
    if (!why.empty()) {
      g_group_error = "context " + std::to_string(k) + ": " + why;
      return PRC_D2Z_ERR_BAD_ARG;
    }
  }
  prc_d2z_group* g = new (std::nothrow) prc_d2z_group();
  if (!g) {
    g_group_error = "host allocation failed";
    return PRC_D2Z_ERR_OOM;
  }
  g->ctx.assign(ctxs, ctxs + n);
  // The boxes are spread over kGroupChains chains (streams) by cells, largest
  // box first onto the least loaded chain; within a chain the boxes of one N
  // form a segment, which runs the whole pass sequence with the kernels
  // compiled for that N.  The chains run side by side, so the latency-bound x
  // pass of one overlaps the copy-bound y pass of another.  Boxes are ordered
  // by (chain, N): every segment is a contiguous range of the descriptors.
  const int nchain = std::min(n, kGroupChains);
  std::vector<int> chain_of(n, 0);
  {
    std::vector<int> order(n);
    for (int k = 0; k < n; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return ctxs[a]->ncell > ctxs[b]->ncell; });
    std::vector<double> load(nchain, 0.0);
    for (int k : order) {
      const int c = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      chain_of[k] = c;
      load[c] += (double)ctxs[k]->ncell;
    }
    std::vector<int> idx(n);
    for (int k = 0; k < n; ++k) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
      return chain_of[a] != chain_of[b] ? chain_of[a] < chain_of[b] : ctxs[a]->N < ctxs[b]->N;
    });
    std::vector<int> ch(n);
    for (int k = 0; k < n; ++k) {
      g->ctx[k] = ctxs[idx[k]];
      ch[k] = chain_of[idx[k]];
    }
    chain_of = ch;
  }
  struct Class {  // a segment: boxes [b0, b1) of one N on one chain
    int b0, b1, N, maxM, zgrid, chain;
  };
  std::vector<Class> classes;
  for (int k = 0; k < n; ++k) {
    if (classes.empty() || classes.back().N != g->ctx[k]->N || classes.back().chain != chain_of[k])
      classes.push_back(Class{k, k, g->ctx[k]->N, 0, 1, chain_of[k]});
    classes.back().b1 = k + 1;
    classes.back().maxM = std::max(classes.back().maxM, g->ctx[k]->wt.M);
  }
  g->device = c0->p.device;
  g->NS = c0->p.Nscales;
  g->nbatch = c0->nbatch;
  g->rb = c0->p.RecordBubbleSizes != 0;
  auto bail = [&](int code, const std::string& m) {
    g_group_error = m;
    prc_d2z_group_destroy(g);
    return code;
  };
  if (hipSetDevice(g->device) != hipSuccess) return bail(PRC_D2Z_ERR_HIP, "hipSetDevice failed");
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(PRC_D2Z_ERR_HIP, "stream creation failed");
  for (int k = 0; k < nchain + 1; ++k) {
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
      return bail(PRC_D2Z_ERR_HIP, "event creation failed");
    g->events.push_back(ev);
  }
  for (int k = 0; k < nchain; ++k) {
    hipStream_t cs = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
      return bail(PRC_D2Z_ERR_HIP, "stream creation failed");
    g->cstreams.push_back(cs);
  }
  for (Class& cl : classes) cl.zgrid = mbox_blocks(MBPass::kFz, g->rb, cl.N);
  std::vector<MBox> boxes(n);
  for (int k = 0; k < n; ++k) {
    prc_d2z_ctx* c = g->ctx[k];
    MBox& B = boxes[k];
    B.delta = c->d_delta;
    B.S = c->d_S;
    B.D0 = c->d_D0;
    B.DN = c->d_DN;
    B.T = c->d_T;
    B.zre = c->d_zre;
    B.rbub = c->p.RecordBubbleSizes ? c->d_rb : nullptr;
    B.rc = c->d_rc;
    B.part = c->d_part;
    B.st_full = c->d_st_full;
    B.st_half = c->d_st_half;
    B.tw_full = c->d_tw_full;
    B.st_fx = c->d_st_fx;
    B.wcoef = c->wt.coef;
    B.sigma_grid = c->d_sigma_grid;
    B.wM = c->wt.M;
    B.winvh = c->wt.invh;
    B.dk = host_dk(c);
    B.invN3 = (float)(1.0 / ((double)c->N * c->N * c->N));
    B.dinit = c->sc.DInit;
    B.N = c->N;
    B.Nscales = c->p.Nscales;
    for (const Class& cl : classes)
      if (cl.N == c->N) B.zgrid = cl.zgrid;
  }
  hipError_t e = hipMalloc(&g->d_boxes, sizeof(MBox) * n);
  if (e == hipSuccess) {
    g->allocs.push_back(g->d_boxes);
    e = hipMemcpy(g->d_boxes, boxes.data(), sizeof(MBox) * n, hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) return bail(PRC_D2Z_ERR_OOM, std::string("group descriptors: ") + hipGetErrorString(e));
  // the launch sequence of one step, per segment, on its chain's stream
  auto add = [&](int c, MBPass pass, int r0, int nb, int toff, int items_nb) -> hipError_t {
    const Class& cl = classes[c];
    prc_d2z_group::Launch L{pass, nullptr, 0, r0, nb, toff, 1, cl.b0, cl.b1 - cl.b0, cl.N, cl.maxM, cl.chain};
    hipError_t ee = group_first(g, pass, items_nb, cl.b0, cl.b1, &L.first, &L.nitems);
    if (ee != hipSuccess) return ee;
    L.nblk = (int)std::max<long long>(1, std::min<long long>(kGroupMaxBlocks, L.nitems));
    if (pass == MBPass::kFz) L.nblk = cl.zgrid;  // the stride of the boxes' partial sums
    g->seq.push_back(L);
    return hipSuccess;
  };
  for (int c = 0; e == hipSuccess && c < (int)classes.size(); ++c) {
    if ((e = add(c, MBPass::kFill, 0, 0, 0, 1)) == hipSuccess && (e = add(c, MBPass::kFwdZ, 0, 0, 0, 1)) == hipSuccess &&
        (e = add(c, MBPass::kFwdY, 0, 0, 0, 1)) == hipSuccess && (e = add(c, MBPass::kFwdX, 0, 0, 0, 1)) == hipSuccess)
      e = add(c, MBPass::kUnpack, 0, 0, 0, 1);
    for (int r0 = 0; e == hipSuccess && r0 < g->NS; r0 += g->nbatch) {
      const int nb = std::min(g->nbatch, g->NS - r0);
      for (int off = 0; e == hipSuccess && off < nb; off += kFxMaxRadii)
        e = add(c, MBPass::kFx, r0 + off, std::min(kFxMaxRadii, nb - off), off, 1);
      if (e == hipSuccess) e = add(c, MBPass::kInvY, r0, nb, 0, nb);
      if (e == hipSuccess) e = add(c, MBPass::kFz, r0, nb, 0, 1);
    }
  }
  if (e != hipSuccess) return bail(PRC_D2Z_ERR_OOM, std::string("group work lists: ") + hipGetErrorString(e));
  *out = g;
  return PRC_D2Z_OK;
}

int prc_d2z_group_run(prc_d2z_group* g) {
  if (!g) return PRC_D2Z_ERR_BAD_ARG;
  (void)hipSetDevice(g->device);
  for (prc_d2z_ctx* c : g->ctx) {
    if (!c->have_delta) return g->fail(PRC_D2Z_ERR_STATE, "a box has no density field");
    if (hipStreamSynchronize(c->stream) != hipSuccess) return g->fail(PRC_D2Z_ERR_HIP, "box stream");
  }
  const int nbox = (int)g->ctx.size();
  // fork the class streams off the group's stream, run each class's chain,
  // join, then the sigma_grid reduction of every box
  auto enqueue = [&]() -> hipError_t {
    const int nc = (int)g->cstreams.size();
    hipError_t e = hipEventRecord(g->events[0], g->stream);
    for (int c = 0; e == hipSuccess && c < nc; ++c) e = hipStreamWaitEvent(g->cstreams[c], g->events[0], 0);
    for (size_t k = 0; e == hipSuccess && k < g->seq.size(); ++k) {
      const auto& L = g->seq[k];
      e = launch_mbox(L.pass, g->d_boxes + L.b0, L.first, L.nbox, L.nitems, L.r0, L.nb, L.toff, L.maxM, L.nblk, L.N,
                      g->rb, g->cstreams[L.cls]);
    }
    for (int c = 0; e == hipSuccess && c < nc; ++c) {
      e = hipEventRecord(g->events[c + 1], g->cstreams[c]);
      if (e == hipSuccess) e = hipStreamWaitEvent(g->stream, g->events[c + 1], 0);
    }
    if (e != hipSuccess) return e;
    return launch_mbox(MBPass::kSigma, g->d_boxes, nullptr, nbox, 0, 0, 0, 0, 0, 1, 0, g->rb, g->stream);
  };
  if (!g->graph) {
    hipError_t e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return g->fail(PRC_D2Z_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    const hipError_t le = enqueue();
    hipGraph_t gr = nullptr;
    e = hipStreamEndCapture(g->stream, &gr);
    if (le != hipSuccess || e != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      return g->fail(PRC_D2Z_ERR_HIP, std::string("group launch: ") + hipGetErrorString(le != hipSuccess ? le : e));
    }
    e = hipGraphInstantiate(&g->graph, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
      g->graph = nullptr;
      return g->fail(PRC_D2Z_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
  }
  hipError_t e = hipGraphLaunch(g->graph, g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) return g->fail(PRC_D2Z_ERR_HIP, std::string("group run: ") + hipGetErrorString(e));
  for (prc_d2z_ctx* c : g->ctx) {
    c->have_spectrum = true;
    c->have_result = true;
  }
  return PRC_D2Z_OK;
}

int prc_d2z_set_graph(prc_d2z_ctx* c, int enable) {
  if (!c) return PRC_D2Z_ERR_BAD_ARG;
  c->use_graph = enable != 0;
  return PRC_D2Z_OK;
}

int prc_d2z_get_timing(prc_d2z_ctx* c, prc_d2z_timing* o) {
  if (!c || !o) return PRC_D2Z_ERR_BAD_ARG;
  memset(o, 0, sizeof(*o));
  if (!c->timer.t0 || c->timer.recs.empty()) return c->fail(PRC_D2Z_ERR_STATE, "no timed run");
  PRC_CHECK(c, hipStreamSynchronize(c->stream), "timing");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c->timer.t0, c->timer.t1);
  o->total_ms = ms;
  const int ncat = std::min((int)c->timer.cats.size(), PRC_D2Z_MAX_KERNELS);
  o->nkernels = ncat;
  for (int k = 0; k < ncat; ++k) strncpy(o->name[k], c->timer.cats[k].c_str(), sizeof(o->name[k]) - 1);
  for (auto& rec : c->timer.recs) {
    if (rec.cat >= ncat) continue;
    float t = 0.f;
    (void)hipEventElapsedTime(&t, rec.a, rec.b);
    o->ms[rec.cat] += t;
    o->launches[rec.cat] += 1;
  }
  return PRC_D2Z_OK;
}

}  // extern "C"
