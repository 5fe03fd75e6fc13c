// Pencil-pass kernels of the delta2zreion path (gfx950).
//
// Layouts (complex elements, C = Blk<N>::C, Nh = N/2, blocked as described in
// d2z_device.h):
//   S   spectrum      [kz/C][ky][kx][kz%C]   written once by the forward passes
//   T_r intermediate  [kz/C][x][y][kz%C]     one per radius of a batch
//   real fields       dense [x][y][z]
// The kz = 0 slot of every row holds the packed pair H(kz=0) + i H(kz=N/2) of
// the z transform (both real along z); the true planes kz = 0 and kz = N/2 of
// the spectrum are kept unpacked in D0, DN [kx][ky].
//
// Passes (reference: main_d2z.C:57-67 forward, delta2zreion_d2z.C:58-146 per
// radius):
//   forward  z r2c (rows -> T layout), y (in place), x (T layout -> S)
//   inverse  x with the filter fused (S -> T_r), y (in place), z c2r with the
//            barrier and running maximum fused (T_r -> zreion)
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

#include "d2z_device.h"
#include "d2z_kernels.h"
#include "fft_core.h"

namespace prc {

// ------------------------------------------------------ intermediate storage
// Intermediates are float2 (FP32 path) or __half2 (mixed path: FP32
// arithmetic, FP16 storage, round to nearest even).
__device__ __forceinline__ float2 ld(const float2* p, long long i) { return p[i]; }
__device__ __forceinline__ float2 ld(const __half2* p, long long i) { return __half22float2(p[i]); }
__device__ __forceinline__ void st(float2* p, long long i, float2 v) { p[i] = v; }
__device__ __forceinline__ void st(__half2* p, long long i, float2 v) { p[i] = __float22half2_rn(v); }

// Offset of element k = i + e*T of z row r in the T layout ([k/C][rows][k%C]).
// When C divides T the offset advances by a constant per e.
template <class P, int C>
__device__ __forceinline__ long long zrow_offset(int i, int e, long long r, long long nrows) {
  constexpr int T = P::T;
  if constexpr (T % C == 0) {
    return ((long long)(i / C) * nrows + r) * C + (i % C) + (long long)e * (T / C) * nrows * C;
  } else {
    const int k = i + e * T;
    return ((long long)(k / C) * nrows + r) * C + (k % C);
  }
}

// ---------------------------------------------------- real-transform glue
// Forward untangling of Z = FFT_L(x[2n] + i x[2n+1]) into the packed half
// spectrum: X[k] = (Z[k] + conj Z[L-k])/2 - i exp(-2 pi i k/N) (Z[k] - conj Z[L-k])/2,
// slot 0 = X[0] + i X[N/2] = (Re Z0 + Im Z0) + i (Re Z0 - Im Z0).
template <class P>
__device__ __forceinline__ void r2c_post(float2 (&v)[P::E], const float2 (&m)[P::E], int i,
                                         const float2* __restrict__ tw_full) {
#pragma unroll
  for (int e = 0; e < P::E; ++e) {
    const int k = i + e * P::T;
    const float2 Z = v[e];
    if (k == 0) {
      v[e] = make_float2(Z.x + Z.y, Z.x - Z.y);
    } else {
      const float2 Zm = cconj(m[e]);
      const float2 Ev = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y + Zm.y));
      const float2 D = make_float2(0.5f * (Z.x - Zm.x), 0.5f * (Z.y - Zm.y));
      const float2 Od = make_float2(D.y, -D.x);  // -i D
      float2 w = tw_full[k];
      w.y = -w.y;
      v[e] = cadd(Ev, cmul(w, Od));
    }
  }
}

// Inverse: Z[k] = X[k] + conj X[L-k] + i exp(+2 pi i k/N) (X[k] - conj X[L-k]),
// Z[0] = (X0 + XN) + i (X0 - XN); then z = IFFT_L(Z) holds x[2n] + i x[2n+1].
template <class P>
__device__ __forceinline__ void c2r_pre(float2 (&v)[P::E], const float2 (&m)[P::E], int i,
                                        const float2* __restrict__ tw_full) {
#pragma unroll
  for (int e = 0; e < P::E; ++e) {
    const int k = i + e * P::T;
    const float2 a = v[e];
    if (k == 0) {
      v[e] = make_float2(a.x + a.y, a.x - a.y);
    } else {
      const float2 b = cconj(m[e]);
      const float2 s = cadd(a, b), d = csub(a, b);
      const float2 wd = cmul(tw_full[k], d);
      v[e] = make_float2(s.x - wd.y, s.y + wd.x);
    }
  }
}

// ------------------------------------------------- forward z pass (r2c)
// Dense real rows (x, y) -> packed half spectrum rows in the T layout.
// LDS (float2) of one 256-thread block of the z-row kernels
template <class P>
constexpr int zrow_lds() {
  return (256 / P::T) * LdsRow<P::L>::STRIDE;
}

// the 256 / T rows from row0 on (one block's work)
template <class P, int C>
__device__ __forceinline__ void fwd_z_rows(const float* __restrict__ in, float2* __restrict__ out, long long nrows,
                                           float dinit, const float2* __restrict__ tw_half,
                                           const float2* __restrict__ tw_full, long long row0, float2* smem) {
  constexpr int L = P::L, T = P::T, E = P::E;
  const int t = threadIdx.x;
  const int i = t % T, rr = t / T;
  const long long row = row0 + rr;
  const bool valid = row < nrows;
  const long long r = valid ? row : 0;
  const float2* src = reinterpret_cast<const float2*>(in + r * (2 * L));
  float2 v[E], m[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float2 a = src[i + e * T];
    v[e] = make_float2(a.x / dinit, a.y / dinit);  // delta /= DInit (main_d2z.C:57-59)
  }
  pencil_fft<P, -1>(v, i, tw_half, LdsRow<L>{smem + rr * LdsRow<L>::STRIDE});
  mirror_lanes<P>(v, m, i);
  r2c_post<P>(v, m, i, tw_full);
  if (valid) {
#pragma unroll
    for (int e = 0; e < E; ++e) out[zrow_offset<P, C>(i, e, row, nrows)] = v[e];
  }
}

template <class P, int C>
__global__ void __launch_bounds__(256) k_fwd_z(const float* __restrict__ in, float2* __restrict__ out,
                                               long long nrows, float dinit, const float2* __restrict__ tw_half,
                                               const float2* __restrict__ tw_full) {
  __shared__ float2 smem[zrow_lds<P>()];
  fwd_z_rows<P, C>(in, out, nrows, dinit, tw_half, tw_full, (long long)blockIdx.x * (256 / P::T), smem);
}

// ------------------------------------------------- tiled complex pass
// One block = one tile of C columns x L points along the transformed axis.
// Tile t -> (o, cb) = (t mod nOuter, t / nOuter); element (a, c) is read at
// cb*sCB + o*in_sO + a*in_sA + c and written at cb*sCB + o*out_sO + a*out_sA + c
// (in place when in == out with equal strides).
template <class P, int C, int DIR, class TI, class TO>
__device__ __forceinline__ void tile_body(const TI* in, TO* out, long long in_sA, long long in_sO, long long out_sA,
                                          long long out_sO, long long sCB, int nOuter, float oscale,
                                          const float2* __restrict__ tw, long long tile, float2* smem) {
  constexpr int T = P::T, E = P::E;
  const int tid = threadIdx.x;
  const int c = tid % C, i = tid / C;
  const long long o = tile % nOuter, cb = tile / nOuter;
  // running pointers: element i + e*T sits e*T*sA after element i
  const TI* ip = in + cb * sCB + o * in_sO + c + (long long)i * in_sA;
  TO* op = out + cb * sCB + o * out_sO + c + (long long)i * out_sA;
  const long long istep = (long long)T * in_sA, ostep = (long long)T * out_sA;
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e, ip += istep) v[e] = ld(ip, 0);
  pencil_fft<P, DIR>(v, i, tw, LdsPencil{smem + c, C});
#pragma unroll
  for (int e = 0; e < E; ++e, op += ostep) st(op, 0, cscale(v[e], oscale));
}

template <class P, int C, int DIR, class TI, class TO>
__global__ void __launch_bounds__(P::T* C)
    k_tile(const TI* in, TO* out, long long in_sA, long long in_sO, long long out_sA, long long out_sO,
           long long sCB, int nOuter, float oscale, const float2* __restrict__ tw) {
  __shared__ float2 smem[P::L * C];
  tile_body<P, C, DIR, TI, TO>(in, out, in_sA, in_sO, out_sA, out_sO, sCB, nOuter, oscale, tw,
                               xcd_tile(blockIdx.x, gridDim.x), smem);
}

// Persistent in-place tile pass (the y passes; TS = float2 or __half2 storage):
// block b walks tiles vb = b + k gridDim.x (through the same XCD-aware map) and
// loads tile vb + gridDim.x into registers, in the storage type, while tile vb
// is transformed and stored, so one tile's loads overlap the previous tile's
// arithmetic.  pencil_fft ends with a barrier after its last LDS read, so the
// LDS is free for the next tile.
template <class P, int C, int DIR, class TS>
__global__ void __launch_bounds__(P::T* C)
    k_tile_ip(TS* data, long long sA, long long sO, long long sCB, int nOuter, long long ntiles, float oscale,
              const float2* __restrict__ tw) {
  constexpr int L = P::L, T = P::T, E = P::E;
  constexpr int NTW = plan_tw_size<P>();
  __shared__ float2 smem[L * C];
  __shared__ float2 s_tw[NTW > 0 ? NTW : 1];  // the plan's twiddles, copied once per (persistent) block
  const int tid = threadIdx.x;
  const int c = tid % C, i = tid / C;
  for (int q = tid; q < NTW; q += T * C) s_tw[q] = tw[q];
  __syncthreads();
  const long long istep = (long long)T * sA;
  const long long lane_off = c + (long long)i * sA;
  auto tile_base = [&](long long vb) {
    const long long tile = xcd_tile(vb, ntiles);
    return (tile / nOuter) * sCB + (tile % nOuter) * sO + lane_off;
  };
  TS raw[E];
  auto fetch = [&](long long vb) {
    const TS* p = data + tile_base(vb);
#pragma unroll
    for (int e = 0; e < E; ++e) raw[e] = p[e * istep];
  };
  long long vb = blockIdx.x;
  if (vb < ntiles) fetch(vb);
  for (; vb < ntiles; vb += gridDim.x) {
    __syncthreads();  // the previous tile's last gather is done (pencil_fft without its tail barrier)
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = ld(raw + e, 0);
    if (vb + gridDim.x < ntiles) fetch(vb + gridDim.x);
    pencil_fft<P, DIR, 0, LdsPencil, false, false>(v, i, s_tw, LdsPencil{smem + c, C});
    TS* op = data + tile_base(vb);
#pragma unroll
    for (int e = 0; e < E; ++e) st(op, e * istep, cscale(v[e], oscale));
  }
}

// ---------------------------------------------------- inverse z pass (c2r)
// Rows (x, y) gathered from the T layout -> dense real rows.
template <class P, int C, class TS>
__global__ void __launch_bounds__(256) k_inv_z(const TS* __restrict__ in, float* __restrict__ out, long long nrows,
                                               float iscale, const float2* __restrict__ tw_half,
                                               const float2* __restrict__ tw_full) {
  constexpr int L = P::L, T = P::T, E = P::E;
  constexpr int ROWS = 256 / T;
  __shared__ float2 smem[ROWS * LdsRow<L>::STRIDE];
  const int t = threadIdx.x;
  const int i = t % T, rr = t / T;
  const long long row = (long long)blockIdx.x * ROWS + rr;
  const bool valid = row < nrows;
  const long long r = valid ? row : 0;
  float2 v[E], m[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = ld(in, zrow_offset<P, C>(i, e, r, nrows));
  mirror_lanes<P>(v, m, i);
  c2r_pre<P>(v, m, i, tw_full);
  pencil_fft<P, +1>(v, i, tw_half, LdsRow<L>{smem + rr * LdsRow<L>::STRIDE});
  if (valid) {
    float2* dst = reinterpret_cast<float2*>(out + row * (2 * L));
#pragma unroll
    for (int e = 0; e < E; ++e) dst[i + e * T] = cscale(v[e], iscale);
  }
}

// ------------------------------------------------- unpack after forward
// D0(k) = (P(k) + conj P(-k))/2,  DN(k) = -i (P(k) - conj P(-k))/2 with P the
// packed kz = 0 slot of S; D0/DN indexed [kx][ky].
template <int C>
__global__ void k_unpack(const float2* __restrict__ S, float2* __restrict__ D0, float2* __restrict__ DN, int N) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long long)N * N) return;
  const int kx = (int)(id / N), ky = (int)(id % N);
  const int mx = (N - kx) % N, my = (N - ky) % N;
  const float2 p = S[((long long)ky * N + kx) * C];
  const float2 q = cconj(S[((long long)my * N + mx) * C]);
  D0[id] = make_float2(0.5f * (p.x + q.x), 0.5f * (p.y + q.y));
  const float2 d = make_float2(0.5f * (p.x - q.x), 0.5f * (p.y - q.y));
  DN[id] = make_float2(d.y, -d.x);
}

// ------------------------------------------------ plain path: filter
// T = W(|k| R)/N^3 * S elementwise; the kz = 0 slot rebuilt as W0 D0 + i WN DN
// (delta2zreion_d2z.C:58-89 with NormalizeArray's 1/N^3, :93).
template <int C>
__global__ void k_filter(const float2* __restrict__ S, const float2* __restrict__ D0, const float2* __restrict__ DN,
                         float2* __restrict__ out, int N, float dkR, float invN3) {
  const long long Nh = N / 2;
  const long long total = (long long)N * N * Nh;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    const SIndex k = s_index<C>(id, N);
    const int mxy = fsq(k.kx, N) + fsq(k.ky, N);
    float2 r;
    if (k.kz != 0) {
      r = cscale(S[id], filter_coef(mxy + k.kz * k.kz, dkR, invN3));
    } else {
      const long long pl = (long long)k.kx * N + k.ky;
      const float w0 = filter_coef(mxy, dkR, invN3);
      const float wn = filter_coef(mxy + (int)(Nh * Nh), dkR, invN3);
      const float2 a = D0[pl], b = DN[pl];
      r = make_float2(w0 * a.x - wn * b.y, w0 * a.y + wn * b.x);
    }
    out[id] = r;
  }
}

// ---------------------------------------- displacement (linear velocity) field
// T = i k_a / |k|^2 * scale * S for component a (0 = x, 1 = y, 2 = z), as
// Delta2Velocity forms xdv_a (delta2velocity_map.C:27-79): k = n dk with the
// signed index n (n >= N/2 -> n - N), the component's own Nyquist plane and
// k = 0 zeroed; the kz = 0 slot is rebuilt as m0 D0 + i mN DN (m = i c).
template <int C>
__global__ void k_disp(const float2* __restrict__ S, const float2* __restrict__ D0, const float2* __restrict__ DN,
                       float2* __restrict__ out, int N, int axis, float dk, float scale) {
  const int Nh = N / 2;
  const long long total = (long long)N * N * Nh;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    const SIndex k = s_index<C>(id, N);
    const int nx = k.kx >= Nh ? k.kx - N : k.kx, ny = k.ky >= Nh ? k.ky - N : k.ky;
    // k_a / |k|^2 * scale at (nx, ny, nz), in the reference's float operations
    auto coef = [&](int nz) -> float {
      const int na = axis == 0 ? nx : (axis == 1 ? ny : nz);
      if (na == Nh || na == -Nh) return 0.0f;
      const float kx = nx * dk, ky = ny * dk, kz = nz * dk;
      const float rk2 = kx * kx + ky * ky + kz * kz;
      if (rk2 == 0.0f) return 0.0f;
      return (na * dk) / rk2 * scale;
    };
    float2 r;
    if (k.kz != 0) {
      const float c = coef(k.kz);
      const float2 a = S[id];
      r = make_float2(-c * a.y, c * a.x);
    } else {
      const long long pl = (long long)k.kx * N + k.ky;
      const float c0 = coef(0), cn = coef(Nh);
      const float2 a = D0[pl], b = DN[pl];
      r = make_float2(-c0 * a.y - cn * b.x, c0 * a.x - cn * b.y);
    }
    out[id] = r;
  }
}

// ------------------------------------------------ binned power spectrum
// PowerSpectrum (powerspectrum_lmb.C:8-139) over the half spectrum kz in
// [0, N/2] (kz = 0 and N/2 from the unpacked planes D0, DN): per mode the
// reference's float k (kx = i dk, minus kmax for i >= N/2), rk2 = kx2 + ky2 +
// kz2, bin = (int)((log(rk) - lnkmin) / dlnk); per bin the count, sum of
// |field|^2 pscale (pscale = DInit^2 / N^3 / d3k: the field is
// FFT(delta) N^-1.5 and S holds FFT(delta / DInit)), sum of 1/rk2 and of
// sqrt(rk2).  Block sums in LDS, then FP64 atomics into out.
template <int C>
__global__ void __launch_bounds__(256)
    k_pk(const float2* __restrict__ S, const float2* __restrict__ D0, const float2* __restrict__ DN, int N, float dk,
         float lnkmin, float dlnk, int nk, double pscale, unsigned long long* __restrict__ count,
         double* __restrict__ sums /* [3][nk]: power, 1/k^2, k */) {
  __shared__ unsigned s_n[kPkMaxBins];
  __shared__ double s_p[kPkMaxBins], s_o[kPkMaxBins], s_k[kPkMaxBins];
  for (int b = threadIdx.x; b < nk; b += blockDim.x) {
    s_n[b] = 0u;
    s_p[b] = s_o[b] = s_k[b] = 0.0;
  }
  __syncthreads();
  const int Nh = N / 2;
  const float kmax = dk * (float)N;
  const long long plane = (long long)N * N, total = plane * (Nh + 1);
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(id / plane), ky = (int)((id / N) % N), kx = (int)(id % N);
    float2 v;
    if (kz == 0)
      v = D0[(long long)kx * N + ky];
    else if (kz == Nh)
      v = DN[(long long)kx * N + ky];
    else
      v = S[(((long long)(kz / C) * N + ky) * N + kx) * C + kz % C];
    float fx = kx * dk, fy = ky * dk, fz = kz * dk;
    if (kx >= Nh) fx = fx - kmax;
    if (ky >= Nh) fy = fy - kmax;
    if (kz >= Nh) fz = fz - kmax;
    const float rk2 = fx * fx + fy * fy + fz * fz;
    const float rk = sqrtf(rk2);
    const float u = (logf(rk) - lnkmin) / dlnk;
    // the reference's int conversion truncates toward zero: u in (-1, 0) (k
    // just below kmin) lands in bin 0; k = 0 gives -inf and no bin
    if (!(u > -1.0f) || u >= (float)nk) continue;
    const int bin = (int)u;
    const double power = ((double)v.x * v.x + (double)v.y * v.y) * pscale;
    atomicAdd(&s_n[bin], 1u);
    atomicAdd(&s_p[bin], power);
    atomicAdd(&s_o[bin], (double)(1.0f / rk2));
    atomicAdd(&s_k[bin], (double)rk);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nk; b += blockDim.x) {
    if (!s_n[b]) continue;
    atomicAdd(&count[b], (unsigned long long)s_n[b]);
    atomicAdd(&sums[b], s_p[b]);
    atomicAdd(&sums[nk + b], s_o[b]);
    atomicAdd(&sums[2 * nk + b], s_k[b]);
  }
}

// ============================================================ fused path
// First inverse pass (x) with the filter fused.  A block owns CF of the C
// columns of one spectrum tile (ky, kz block); sibling sub-tiles are adjacent
// in dispatch order, so their partial lines meet in the same L2.  The
// spectrum values are read once into registers and reused for nb radii: each
// radius scales by W(|k| R_r)/N^3, transforms along x and writes its own
// intermediate T_r = Tbuf + r_local * N*N*N/2 (T layout).
//
// W depends on kx only through |fx| (fx = kx or kx - N), so per radius the
// block first tabulates W in LDS for its CF columns and the N/2 + 1 values of
// |fx| (half the filter evaluations of a per-element loop), plus the kz = N/2
// column for the packed slot.  |k| of every table entry is tabulated once
// before the radius loop.  The table is laid out [|fx|][column] so that the
// lanes of a wave (consecutive columns, then consecutive kx) read consecutive
// words.
// LDS of one fused x block: the padded exchange buffer (float2) and the |k|
// table (float)
template <class P, int CF>
struct FxLds {
  static constexpr int SH = ilog2(P::radices[0]);
  static constexpr int NROW = P::L + (P::L >> SH);  // padded exchange rows
  static constexpr int SMEM = NROW * CF;           // float2
  static constexpr int KTAB = (CF + 1) * (P::L / 2 + 1);  // float
};

// One tile (ky, kz block, column sub-tile) of the fused x pass for nb radii;
// wsh (the W(x) table, wM intervals) must be in LDS before the call.
template <class P, int C, int CF, class TS>
__device__ __forceinline__ void fx_body(const float2* __restrict__ S, const float2* __restrict__ D0,
                                        const float2* __restrict__ DN, TS* __restrict__ Tbuf, float dk, float invN3,
                                        const RadiusConst* __restrict__ rc, int r0, int nb,
                                        const float2* __restrict__ tw, const float4* wsh, int wM, float winvh,
                                        long long tile, float2* smem, float* ktab) {
  constexpr int N = P::L, T = P::T, E = P::E;
  constexpr int NSUB = C / CF;
  constexpr int NF = N / 2 + 1;           // distinct |fx|
  constexpr int NT = P::T * CF;           // threads
  constexpr long long Nh = N / 2;
  constexpr long long plane = (long long)N * N * Nh;
  constexpr int SH = FxLds<P, CF>::SH;
  constexpr int NROW = FxLds<P, CF>::NROW;
  // W table [|fx|][CF] then the kz = N/2 column [|fx|]; it aliases the
  // transform's exchange buffer, which is idle between radii
  static_assert((CF + 1) * NF <= 2 * NROW * CF, "W table must fit in the exchange buffer");
  float* wtab = reinterpret_cast<float*>(smem);
  float* wpk = wtab + CF * NF;
  const int tid = threadIdx.x;
  const int cl = tid % CF, i = tid / CF;
  const int sub = (int)(tile % NSUB);
  const int ky = (int)((tile / NSUB) % N);
  const int kzb = (int)(tile / (NSUB * N));
  const int kz0 = kzb * C + sub * CF;
  const int kz = kz0 + cl;
  const bool packed = (kz == 0);
  const bool has_packed = (kzb == 0 && sub == 0);  // block-uniform
  const float2* sp = S + ((long long)kzb * N + ky) * N * C + sub * CF + cl;  // element kx at sp[kx*C]
  float2 m[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kx = i + e * T;
    m[e] = packed ? D0[(long long)kx * N + ky] : sp[(long long)kx * C];
  }
  const int fy2 = fsq(ky, N);
  TS* tp = Tbuf + (long long)kzb * N * N * C + (long long)ky * C + sub * CF + cl;  // element x at tp[x*N*C]
  const LdsColPad<SH> lds{smem + cl, CF};
  const int ntab = (has_packed ? CF + 1 : CF) * NF;
  for (int q = tid; q < ntab; q += NT) {
    int f, kzq;
    if (q < CF * NF) {
      f = q / CF;
      kzq = kz0 + q % CF;
    } else {
      f = q - CF * NF;
      kzq = N / 2;
    }
    ktab[q] = kmag_grid(f * f + fy2 + kzq * kzq);
  }
  __syncthreads();  // ktab and the W(x) table complete
#pragma unroll 1
  for (int rl = 0; rl < nb; ++rl) {
    __syncthreads();  // the last radius's final gather is done (pencil_fft without its tail barrier)
    const float dkR = dk * rc[r0 + rl].R;
    const float dkRh = dkR * winvh;
    // tabulate W(|k| R)/N^3 for this block's columns (the exchange buffer is
    // free after the barrier above)
    fill_wtable<NT, ((CF + 1) * NF + NT - 1) / NT>(wtab, ktab, ntab, dkR, dkRh, invN3, wsh, wM, tid);
    __syncthreads();
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int kx = i + e * T;
      const int fa = (kx <= N / 2) ? kx : N - kx;
      const float w0 = wtab[fa * CF + cl];
      if (!packed) {
        v[e] = cscale(m[e], w0);
      } else {
        const float wn = wpk[fa];
        const float2 b = DN[(long long)kx * N + ky];
        v[e] = make_float2(w0 * m[e].x - wn * b.y, w0 * m[e].y + wn * b.x);
      }
    }
    lds_hold(v);      // the table reads stay above the barrier
    __syncthreads();  // table reads done before the transform reuses the buffer
    pencil_fft<P, +1, 0, LdsColPad<SH>, false, false>(v, i, tw, lds);
    TS* dst = tp + (long long)rl * plane;
#pragma unroll
    for (int e = 0; e < E; ++e) st(dst, (long long)(i + e * T) * N * C, v[e]);
  }
}

template <class P, int C, int CF, class TS>
__global__ void __launch_bounds__(P::T* CF)
    k_fx(const float2* __restrict__ S, const float2* __restrict__ D0, const float2* __restrict__ DN,
         TS* __restrict__ Tbuf, float dk, float invN3, const RadiusConst* __restrict__ rc, int r0, int nb,
         const float2* __restrict__ tw, const float4* __restrict__ wcoef, int wM, float winvh) {
  __shared__ float2 smem[FxLds<P, CF>::SMEM];
  __shared__ float ktab[FxLds<P, CF>::KTAB];  // |k|/dk of the table entries, same order
  extern __shared__ float4 wsh[];              // W(x) interpolation table (wM intervals)
  for (int q = threadIdx.x; q < wM; q += P::T * CF) wsh[q] = wcoef[q];
  // FP32 outputs: consecutive lines over the XCDs; FP16: one range per XCD
  const long long tile = std::is_same<TS, float2>::value
                             ? line_tile(blockIdx.x, gridDim.x, line_group(C / CF, C * (int)sizeof(TS)))
                             : xcd_tile(blockIdx.x, gridDim.x);
  fx_body<P, C, CF, TS>(S, D0, DN, Tbuf, dk, invN3, rc, r0, nb, tw, wsh, wM, winvh, tile, smem, ktab);
}

// columns per fused-x block: keep the block at <= 512 threads
// Plan of the fused x pass.  It keeps a second register copy of its tile (the
// unfiltered spectrum, reused for every radius), so at N = 1024 and 2048 it
// uses 8 elements per thread (1024 = 8*8*4*4, 2048 = 8*8*8*4; 128 VGPRs)
// instead of the plain passes' 16 (221-225 VGPRs): 16 waves per CU instead
// of 8.
template <int N>
struct FxPlan {
  using P = TilePlan<N>;
};
template <>
struct FxPlan<1024> {
  using P = PlanT<1024, 8, 8, 8, 4, 4>;
};
template <>
struct FxPlan<2048> {
  using P = PlanT<2048, 8, 8, 8, 8, 4>;
};

// Columns per fused x block: at most 512 threads, except with the 8-element
// plans, which take 1024 threads: 8 columns x 128 at N = 1024 (64 B segments
// per row; 4 columns, 32 B segments, made the FP32 stores far slower) and
// 4 columns x 256 at N = 2048 (C = 8 there).
template <int N>
struct FxCols {
  using P = typename FxPlan<N>::P;
  static constexpr int C = Blk<N>::C;
  static constexpr int CF = (N == 1024)   ? 8
                            : (N == 2048) ? 4
                                          : ((P::T * C <= 512) ? C : (P::T * (C / 2) <= 512 ? C / 2 : C / 4));
};

// Last inverse pass (z, c2r) with the barrier fused: for each row the nb
// intermediates are transformed in turn, d2z() is evaluated per cell, the
// maximum over the batch stays in registers (first radius wins ties) and
// zreion (and rbubble) is read and written once per batch.  Per-radius sums of
// delta_R^2 (float products, double sums as GetMeanSigma; with FP16 storage
// the wave's own sum is float) are kept per wave in
// LDS and written as one partial per block and radius.
// LDS of one fused z block: the pencil rows, the per-wave sums, the plan's
// twiddles (transposed) and the c2r glue constants
template <class P>
struct FzLds {
  static constexpr int NTW = plan_tw_size<P>() > 0 ? plan_tw_size<P>() : 1;
  static constexpr int SMEM = (256 / P::T) * LdsRow<P::L>::STRIDE;  // float2
};

// Fills the z block's LDS tables: red = 0, the plan's twiddles transposed, w^k
// (FP32 storage) or the packed FP16 glue constants (FP16 storage).  A barrier
// must follow before fz_group reads them.
template <class P, class TS>
__device__ __forceinline__ void fz_tables(double (*red)[kFusedMaxRadii], float2* s_twh, float2* s_twf,
                                          const float2* __restrict__ tw_half, const float2* __restrict__ tw_full) {
  constexpr int L = P::L;
  constexpr bool kH = std::is_same<TS, __half2>::value;
  const int t = threadIdx.x;
  if (t < 4 * kFusedMaxRadii) red[t / kFusedMaxRadii][t % kFusedMaxRadii] = 0.0;
  copy_twiddles_transposed<P>(s_twh, tw_half, t, 256);
  uint2* const s_pq = reinterpret_cast<uint2*>(s_twf);
  for (int q = t; q < L; q += 256) {
    if constexpr (kH)
      s_pq[q] = c2r_glue_pq(q, tw_full[q]);
    else
      s_twf[q] = tw_full[q];
  }
}

// One group of 256 / T z rows from base on, for the nb radii of a batch: each
// row's intermediates are transformed in turn, d2z() is evaluated per cell,
// the maximum over the batch stays in registers (first radius wins ties) and
// zreion (and rbubble) is read and written once.  Per-radius sums of
// delta_R^2 (float products, double sums as GetMeanSigma; with FP16 storage
// the wave's own sum is float) are added per wave into red.  F32B: the
// barrier as one FP32 FMA (barrier_z_f32), as for FP16 storage, instead of the
// reference's double evaluation (the mixed path's FP32 intermediates).
template <class P, int C, bool RB, class TS, bool F32B = false>
__device__ __forceinline__ void fz_group(const TS* __restrict__ Tbuf, float* __restrict__ zre,
                                         float* __restrict__ rbub, long long nrows,
                                         const RadiusConst* __restrict__ rc, int r0, int nb, float iscale,
                                         long long base, const float2* s_twh, const float2* s_twf, float2* smem,
                                         double (*red)[kFusedMaxRadii]) {
  constexpr int L = P::L, T = P::T, E = P::E;
  constexpr bool kH = std::is_same<TS, __half2>::value;
  const uint2* const s_pq = reinterpret_cast<const uint2*>(s_twf);
  const int t = threadIdx.x;
  const int i = t % T, rr = t / T;
  const int w = t >> 6, lane = t & 63;
  const long long plane = nrows * L;
  // FP16 storage with 8-element pencils (N <= 1024): the pass is issue-bound
  // and each wave synchronises its own pencil.  Otherwise block barriers keep
  // the waves' loads (adjacent segments of each kz block) together in time:
  // wave-level sync measured 8% slower for FP32 at N = 1024 and 40% slower
  // for FP16 at N = 2048.
  using Lds = LdsRow<L, kH && (E <= 8)>;
  const Lds lds{smem + rr * Lds::STRIDE};
  const double sc = (double)iscale, sc2 = sc * sc;
  const long long row = base + rr;
  const bool valid = row < nrows;
  const long long r = valid ? row : 0;
  float2 best[E];
  float bR0[E], bR1[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    best[e] = make_float2(0.0f, 0.0f);  // zreion >= 0: see zmax_nonneg
    bR0[e] = 0.f;
    bR1[e] = 0.f;
  }
  // FP32 storage at N = 1024 (8-element pencils), no bubble radii: the next
  // radius's row is loaded while this one is transformed (16 more VGPRs, 112
  // -> 124, still 4 waves per SIMD; z pass 110 -> 107.5 ms per 1024^3 x 96
  // step, same bits).  The other variants would lose a wave per SIMD.
  constexpr bool kPre = !kH && !RB && E <= 8 && L == 512;
  float2 nx[kPre ? E : 1];
  if constexpr (kPre) {
#pragma unroll
    for (int e = 0; e < E; ++e) nx[e] = ld(Tbuf, zrow_offset<P, C>(i, e, r, nrows));
  }
#pragma unroll 1
  for (int q = 0; q < nb; ++q) {
    const TS* tp = Tbuf + (long long)q * plane;
    float2 v[E];
    if constexpr (kPre) {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = nx[e];
      const TS* tn = Tbuf + (long long)min(q + 1, nb - 1) * plane;  // the last radius reloads its own row
#pragma unroll
      for (int e = 0; e < E; ++e) nx[e] = ld(tn, zrow_offset<P, C>(i, e, r, nrows));
    }
    // read per radius from LDS rather than held in registers across radii
    int zero = 0;
    asm volatile("" : "+s"(zero));
    if constexpr (kH) {  // glue on the stored FP16 pairs: one move and four packed FMAs each
      const uint32_t* tu = reinterpret_cast<const uint32_t*>(tp);
      uint32_t a[E], m[E];
#pragma unroll
      for (int e = 0; e < E; ++e) a[e] = tu[zrow_offset<P, C>(i, e, r, nrows)];
      mirror_lanes<P>(a, m, i);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const uint2 pq = s_pq[zero + i + e * T];
        v[e] = unpack_h2(c2r_glue_h2(a[e], m[e], pq.x, pq.y));
      }
    } else {
      float2 m[E];
      if constexpr (!kPre) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = ld(tp, zrow_offset<P, C>(i, e, r, nrows));
      }
      mirror_lanes<P>(v, m, i);
      c2r_pre<P>(v, m, i, s_twf + zero);
    }
    pencil_fft<P, +1, 0, Lds, true>(v, i, s_twh + zero, lds);
    // the output scale iscale (a power of two) is applied inside the
    // barrier's FMA and to the sum of squares, where it is exact (the FP16
    // storage and F32B evaluate it in FP32: |error| <= (|z| + |A B - 1|) 2^-24)
    const double A = rc[r0 + q].A, B = rc[r0 + q].B;
    const BarrierF32 bk = barrier_f32(A, B, iscale);
    const float R = rc[r0 + q].R;
    float acc = 0.0f;  // this thread's 2E cells; blocks and waves are summed in double
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const float z0 = (kH || F32B) ? barrier_z_f32(v[e].x, bk) : barrier_z_scaled(v[e].x, sc, A, B);
      const float z1 = (kH || F32B) ? barrier_z_f32(v[e].y, bk) : barrier_z_scaled(v[e].y, sc, A, B);
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
      } else {  // same zreion as the strict comparison: z is never NaN (a radius
                // with sigma(R) > sigma_min has B = -inf, see barrier_offset)
        best[e].x = zmax_nonneg(best[e].x, z0);
        best[e].y = zmax_nonneg(best[e].y, z1);
      }
    }
    const double wsum = kH ? (double)wave_sum_f(valid ? acc : 0.0f) * sc2 : wave_sum(valid ? (double)acc * sc2 : 0.0);
    if (lane == 0) red[w][q] += wsum;
  }
  if (valid) {
    float2* zp = reinterpret_cast<float2*>(zre + row * (2 * L));
    float2* bp = RB ? reinterpret_cast<float2*>(rbub + row * (2 * L)) : nullptr;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int n = i + e * T;
      const float2 old = zp[n];
      const bool u0 = best[e].x > old.x, u1 = best[e].y > old.y;
      if (u0 || u1) {
        zp[n] = make_float2(u0 ? best[e].x : old.x, u1 ? best[e].y : old.y);
        if (RB) {
          const float2 ob = bp[n];
          bp[n] = make_float2(u0 ? bR0[e] : ob.x, u1 ? bR1[e] : ob.y);
        }
      }
    }
  }
}

// The fused z pass (c2r + barrier + running max) over all rows: a
// grid-stride kernel over row groups; one partial sum per block and radius.
template <class P, int C, bool RB, class TS, bool F32B = false>
__global__ void __launch_bounds__(256)
    k_fz(const TS* __restrict__ Tbuf, float* __restrict__ zre, float* __restrict__ rbub, long long nrows,
         const RadiusConst* __restrict__ rc, int r0, int nb, double* __restrict__ partial, float iscale,
         const float2* __restrict__ tw_half, const float2* __restrict__ tw_full) {
  constexpr int ROWS = 256 / P::T;
  __shared__ float2 smem[FzLds<P>::SMEM];
  __shared__ double red[4][kFusedMaxRadii];
  // twiddles in LDS (the blocks are persistent: copied once), the plan's
  // stages transposed so that lanes with consecutive k read consecutive words;
  // c2r glue: w^k for FP32 storage, the packed FP16 constants (P, Q) of
  // c2r_glue_h2 for FP16 storage
  __shared__ float2 s_twh[FzLds<P>::NTW];
  __shared__ float2 s_twf[P::L];
  fz_tables<P, TS>(red, s_twh, s_twf, tw_half, tw_full);
  __syncthreads();
#pragma unroll 1
  for (long long base = (long long)blockIdx.x * ROWS; base < nrows; base += (long long)gridDim.x * ROWS)
    fz_group<P, C, RB, TS, F32B>(Tbuf, zre, rbub, nrows, rc, r0, nb, iscale, base, s_twh, s_twf, smem, red);
  __syncthreads();
  const int t = threadIdx.x;
  if (t < nb) {
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k][t];
    partial[(long long)(r0 + t) * gridDim.x + blockIdx.x] = s;
  }
}

// ============================================= multi-box persistent kernels
// One launch per pass covers every box of a group (N in {64, 128, 192, 256});
// grid = the kernel's resident capacity, and each block walks the work list:
// item t of round k is t = k G + (block's slot), the slot XCD-aware inside
// full rounds (consecutive items on one XCD, as the single-box kernels'
// tiles).  Box b owns items [first[b], first[b+1]).  Every item runs the body
// of the single-box kernel for its N, on that box's buffers.
template <class F>
__device__ __forceinline__ void mb_switch(int N, F&& f) {
  switch (N) {
    case 64: f(std::integral_constant<int, 64>{}); break;
    case 128: f(std::integral_constant<int, 128>{}); break;
    case 192: f(std::integral_constant<int, 192>{}); break;
    case 256: f(std::integral_constant<int, 256>{}); break;
    default: break;
  }
}

// the same dispatch for host code (launch shapes of the per-size kernels)
template <class F>
inline void mb_switch_host(int N, F&& f) {
  switch (N) {
    case 64: f(std::integral_constant<int, 64>{}); break;
    case 128: f(std::integral_constant<int, 128>{}); break;
    case 192: f(std::integral_constant<int, 192>{}); break;
    case 256: f(std::integral_constant<int, 256>{}); break;
    default: break;
  }
}

__device__ __forceinline__ int mbox_of(const long long* __restrict__ first, int nbox, long long t) {
  int lo = 0, hi = nbox - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// item of round k for this block, or -1 past the end of the list
__device__ __forceinline__ long long mbox_item(long long k, long long nitems) {
  const long long G = gridDim.x, base = k * G;
  if (base + G <= nitems) return base + xcd_tile(blockIdx.x, G);
  return (base + blockIdx.x < nitems) ? base + blockIdx.x : -1;
}

template <int N>
__device__ __forceinline__ TileMap mb_tile_map(MBPass pass, int nb) {
  constexpr int C = Blk<N>::C;
  constexpr long long NC = (long long)N * C;
  TileMap m{};
  m.nOuter = N;
  m.nColBlocks = (N / 2) / C;
  m.sCB = (long long)N * N * C;
  if (pass == MBPass::kFwdX) {  // T layout (x, ky) -> S layout (ky, kx)
    m.in_sA = NC;
    m.in_sO = C;
    m.out_sA = C;
    m.out_sO = NC;
  } else {  // y in place; the inverse y covers the batch's nb intermediates
    m.in_sA = m.out_sA = C;
    m.in_sO = m.out_sO = NC;
    if (pass == MBPass::kInvY) m.nColBlocks *= nb;
  }
  return m;
}

template <int N>
constexpr long long mb_items(MBPass pass, int nb) {
  constexpr long long n2 = (long long)N * N;
  switch (pass) {
    case MBPass::kFill: return (n2 * N + 4095) / 4096;
    case MBPass::kFwdZ: return n2 / (256 / Plan<N / 2>::T);
    case MBPass::kFwdY:
    case MBPass::kFwdX: return (long long)N * ((N / 2) / Blk<N>::C);
    case MBPass::kInvY: return (long long)N * ((N / 2) / Blk<N>::C) * nb;
    case MBPass::kUnpack: return (n2 + 255) / 256;
    case MBPass::kFx: return (long long)N * ((N / 2) / FxCols<N>::C) * (FxCols<N>::C / FxCols<N>::CF);
    case MBPass::kFz: return (n2 + (256 / Plan<N / 2>::T) - 1) / (256 / Plan<N / 2>::T);
    default: return 0;
  }
}

// LDS (bytes) of each pass's body, the largest over the supported N
template <int N>
constexpr size_t mb_lds_n(MBPass pass) {
  using PZ = Plan<N / 2>;
  using PF = typename FxCols<N>::P;
  switch (pass) {
    case MBPass::kFwdZ: return sizeof(float2) * zrow_lds<PZ>();
    case MBPass::kFwdY:
    case MBPass::kFwdX:
    case MBPass::kInvY: return sizeof(float2) * TilePlan<N>::L * Blk<N>::C;
    case MBPass::kFx:
      return sizeof(float2) * FxLds<PF, FxCols<N>::CF>::SMEM + sizeof(float) * FxLds<PF, FxCols<N>::CF>::KTAB;
    case MBPass::kFz:
      return sizeof(float2) * (FzLds<PZ>::SMEM + FzLds<PZ>::NTW + PZ::L);
    default: return 0;
  }
}
constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
constexpr size_t mb_lds(MBPass pass) {
  return cmax(cmax(mb_lds_n<64>(pass), mb_lds_n<128>(pass)), cmax(mb_lds_n<192>(pass), mb_lds_n<256>(pass)));
}
__global__ void __launch_bounds__(256)
    k_mb_fill(const MBox* __restrict__ boxes, const long long* __restrict__ first, int nbox, long long nitems) {
  for (long long k = 0;; ++k) {
    const long long t = mbox_item(k, nitems);
    if (t < 0) break;
    const int b = mbox_of(first, nbox, t);
    const long long n = (long long)boxes[b].N * boxes[b].N * boxes[b].N;
    const long long i0 = (t - first[b]) * 4096;
    float* zre = boxes[b].zre;
    float* rb = boxes[b].rbub;
    for (long long i = i0 + threadIdx.x; i < i0 + 4096 && i < n; i += 256) {
      zre[i] = 0.0f;
      if (rb) rb[i] = 0.0f;
    }
  }
}

__global__ void __launch_bounds__(256)
    k_mb_fwdz(const MBox* __restrict__ boxes, const long long* __restrict__ first, int nbox, long long nitems) {
  extern __shared__ float4 mb_lds4[];
  float2* smem = reinterpret_cast<float2*>(mb_lds4);
  for (long long k = 0;; ++k) {
    const long long t = mbox_item(k, nitems);
    if (t < 0) break;
    const int b = mbox_of(first, nbox, t);
    const MBox& B = boxes[b];
    const long long item = t - first[b];
    mb_switch(B.N, [&](auto nc) {
      constexpr int N = decltype(nc)::value;
      using P = Plan<N / 2>;
      fwd_z_rows<P, Blk<N>::C>(B.delta, B.T, (long long)N * N, B.dinit, B.st_half, B.tw_full, item * (256 / P::T),
                               smem);
    });
    __syncthreads();  // the LDS is reused by the next item
  }
}

template <int DIR>
__global__ void __launch_bounds__(256) k_mb_tile(const MBox* __restrict__ boxes, const long long* __restrict__ first,
                                                 int nbox, long long nitems, MBPass pass, int nb) {
  extern __shared__ float4 mb_lds4[];
  float2* smem = reinterpret_cast<float2*>(mb_lds4);
  for (long long k = 0;; ++k) {
    const long long t = mbox_item(k, nitems);
    if (t < 0) break;
    const int b = mbox_of(first, nbox, t);
    const MBox& B = boxes[b];
    const long long item = t - first[b];
    mb_switch(B.N, [&](auto nc) {
      constexpr int N = decltype(nc)::value;
      using P = TilePlan<N>;
      static_assert(P::T * Blk<N>::C == 256, "multi-box tiles are 256-thread blocks");
      const TileMap m = mb_tile_map<N>(pass, nb);
      const float2* in = B.T;
      float2* out = (pass == MBPass::kFwdX) ? B.S : B.T;
      tile_body<P, Blk<N>::C, DIR, float2, float2>(in, out, m.in_sA, m.in_sO, m.out_sA, m.out_sO, m.sCB, m.nOuter,
                                                   1.0f, B.st_full, item, smem);
    });
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256)
    k_mb_unpack(const MBox* __restrict__ boxes, const long long* __restrict__ first, int nbox, long long nitems) {
  for (long long k = 0;; ++k) {
    const long long t = mbox_item(k, nitems);
    if (t < 0) break;
    const int b = mbox_of(first, nbox, t);
    const MBox& B = boxes[b];
    const int N = B.N;
    const long long id = (t - first[b]) * 256 + threadIdx.x;
    if (id >= (long long)N * N) continue;
    const int kx = (int)(id / N), ky = (int)(id % N);
    const int mx = (N - kx) % N, my = (N - ky) % N;
    constexpr int C = 16;  // Blk<N>::C for every multi-box N
    const float2 p = B.S[((long long)ky * N + kx) * C];
    const float2 q = cconj(B.S[((long long)my * N + mx) * C]);
    B.D0[id] = make_float2(0.5f * (p.x + q.x), 0.5f * (p.y + q.y));
    const float2 d = make_float2(0.5f * (p.x - q.x), 0.5f * (p.y - q.y));
    B.DN[id] = make_float2(d.y, -d.x);
  }
}

// the x pass of one size class: the box's W(x) table is loaded into LDS when
// the block moves to a new box (block-uniform); the body's first barrier
// covers the load
template <int N>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    k_mb_fx(const MBox* __restrict__ boxes, const long long* __restrict__ first,
                                               int nbox, long long nitems, int r0, int nb, int toff, int maxM) {
  using P = typename FxCols<N>::P;
  constexpr int C = FxCols<N>::C, CF = FxCols<N>::CF;
  static_assert(P::T * CF == 256, "multi-box x tiles are 256-thread blocks");
  extern __shared__ float4 mb_lds4[];
  float4* wsh = mb_lds4;
  float2* smem = reinterpret_cast<float2*>(mb_lds4 + maxM);
  float* ktab = reinterpret_cast<float*>(smem + FxLds<P, CF>::SMEM);
  int loaded = -1;
  for (long long k = 0;; ++k) {
    const long long t = mbox_item(k, nitems);
    if (t < 0) break;
    const int b = mbox_of(first, nbox, t);
    const MBox& B = boxes[b];
    if (b != loaded) {
      __syncthreads();  // the previous box's table is no longer read
      for (int q = threadIdx.x; q < B.wM; q += 256) wsh[q] = B.wcoef[q];
      loaded = b;
    }
    float2* T = B.T + (long long)toff * ((long long)N * N * (N / 2));
    fx_body<P, C, CF, float2>(B.S, B.D0, B.DN, T, B.dk, B.invN3, B.rc, r0, nb, B.st_fx, wsh, B.wM, B.winvh,
                              t - first[b], smem, ktab);
    __syncthreads();
  }
}

// the z pass of one size class: per-radius sums of squares of the box a block
// is on are kept in LDS and written to that box's partials (slot = block) when
// the block moves on; each block first zeroes its slot of every box, so boxes
// it never visits read 0 there
template <int N, bool RB>
__global__ void __launch_bounds__(256) k_mb_fz(const MBox* __restrict__ boxes, const long long* __restrict__ first,
                                               int nbox, long long nitems, int r0, int nb) {
  using P = Plan<N / 2>;
  __shared__ float2 smem[FzLds<P>::SMEM];
  __shared__ float2 s_twh[FzLds<P>::NTW];
  __shared__ float2 s_twf[P::L];
  __shared__ double red[4][kFusedMaxRadii];
  const long long G = gridDim.x;
  for (int q = threadIdx.x; q < nbox * nb; q += 256)
    boxes[q / nb].part[(long long)(r0 + q % nb) * G + blockIdx.x] = 0.0;
  auto flush = [&](int b) {
    __syncthreads();
    if ((int)threadIdx.x < nb) {
      double s = 0.0;
      for (int w = 0; w < 4; ++w) s += red[w][threadIdx.x];
      boxes[b].part[(long long)(r0 + threadIdx.x) * G + blockIdx.x] = s;
    }
    __syncthreads();
  };
  int cur = -1;
  for (long long k = 0;; ++k) {
    const long long t = mbox_item(k, nitems);
    if (t < 0) break;
    const int b = mbox_of(first, nbox, t);
    const MBox& B = boxes[b];
    if (b != cur) {
      if (cur >= 0) flush(cur);
      else __syncthreads();  // the zeroing above precedes this block's first flush
      fz_tables<P, float2>(red, s_twh, s_twf, B.st_half, B.tw_full);
      __syncthreads();
      cur = b;
    }
    fz_group<P, Blk<N>::C, RB, float2>(B.T, B.zre, RB ? B.rbub : nullptr, (long long)N * N, B.rc, r0, nb, 1.0f,
                                       (t - first[b]) * (256 / P::T), s_twh, s_twf, smem, red);
  }
  if (cur >= 0) flush(cur);
}

// sigma_grid of every box and radius from the z pass's partials (the box's
// class grid of slots)
__global__ void k_mb_sigma(const MBox* __restrict__ boxes, int nbox) {
  const int b = blockIdx.x / kFusedMaxRadii / 8, r = blockIdx.x % (kFusedMaxRadii * 8);
  if (b >= nbox || r >= boxes[b].Nscales) return;
  const MBox& B = boxes[b];
  double a = 0.0;
  for (int k = threadIdx.x; k < B.zgrid; k += 64) a += B.part[(long long)r * B.zgrid + k];
  a = wave_sum(a);
  if (threadIdx.x == 0) B.sigma_grid[r] = (float)sqrt(a / ((double)B.N * B.N * B.N));
}

bool mbox_supported(int N) { return N == 64 || N == 128 || N == 192 || N == 256; }

long long mbox_items(MBPass pass, int N, int nb) {
  switch (N) {
    case 64: return mb_items<64>(pass, nb);
    case 128: return mb_items<128>(pass, nb);
    case 192: return mb_items<192>(pass, nb);
    case 256: return mb_items<256>(pass, nb);
    default: return 0;
  }
}

template <class F>
static int mb_resident(F f, size_t lds) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 256, lds);
  int g = (e == hipSuccess && cus > 0 && per_cu > 0) ? cus * per_cu : 1024;
  return g > kFzMaxBlocks ? kFzMaxBlocks : g;
}

// resident capacity of the x and z kernels of size class N
template <int N>
static int mb_blocks_n(MBPass pass, bool rb) {
  if (pass == MBPass::kFx) return mb_resident(k_mb_fx<N>, mb_lds_n<N>(pass) + sizeof(float4) * kWTableMax);
  return rb ? mb_resident(k_mb_fz<N, true>, 0) : mb_resident(k_mb_fz<N, false>, 0);
}

int mbox_blocks(MBPass pass, bool rb, int N) {
  switch (pass) {
    case MBPass::kFill: return mb_resident(k_mb_fill, 0);
    case MBPass::kFwdZ: return mb_resident(k_mb_fwdz, mb_lds(pass));
    case MBPass::kFwdY:
    case MBPass::kFwdX: return mb_resident(k_mb_tile<-1>, mb_lds(pass));
    case MBPass::kInvY: return mb_resident(k_mb_tile<+1>, mb_lds(pass));
    case MBPass::kUnpack: return mb_resident(k_mb_unpack, 0);
    case MBPass::kFx:
    case MBPass::kFz: {
      int g = 1;
      mb_switch_host(N, [&](auto nc) { g = mb_blocks_n<decltype(nc)::value>(pass, rb); });
      return g;
    }
    default: return 1;
  }
}

template <int N>
static void launch_mb_class(MBPass pass, const MBox* boxes, const long long* first, int nbox, long long nitems, int r0,
                            int nb, int toff, int maxM, dim3 g, bool rb, hipStream_t s) {
  const dim3 b(256);
  if (pass == MBPass::kFx)
    hipLaunchKernelGGL(k_mb_fx<N>, g, b, mb_lds_n<N>(pass) + sizeof(float4) * maxM, s, boxes, first, nbox, nitems,
                       r0, nb, toff, maxM);
  else if (rb)
    hipLaunchKernelGGL((k_mb_fz<N, true>), g, b, 0, s, boxes, first, nbox, nitems, r0, nb);
  else
    hipLaunchKernelGGL((k_mb_fz<N, false>), g, b, 0, s, boxes, first, nbox, nitems, r0, nb);
}

hipError_t launch_mbox(MBPass pass, const MBox* boxes, const long long* first, int nbox, long long nitems, int r0,
                       int nb, int toff, int maxM, int nblk, int N, bool rb, hipStream_t s) {
  if (nbox < 1 || nbox > kMBoxMax || nblk < 1 || nitems < 0) return hipErrorInvalidValue;
  const dim3 g(nblk), b(256);
  switch (pass) {
    case MBPass::kFill:
      hipLaunchKernelGGL(k_mb_fill, g, b, 0, s, boxes, first, nbox, nitems);
      break;
    case MBPass::kFwdZ:
      hipLaunchKernelGGL(k_mb_fwdz, g, b, mb_lds(pass), s, boxes, first, nbox, nitems);
      break;
    case MBPass::kFwdY:
    case MBPass::kFwdX:
      hipLaunchKernelGGL(k_mb_tile<-1>, g, b, mb_lds(pass), s, boxes, first, nbox, nitems, pass, nb);
      break;
    case MBPass::kInvY:
      hipLaunchKernelGGL(k_mb_tile<+1>, g, b, mb_lds(pass), s, boxes, first, nbox, nitems, pass, nb);
      break;
    case MBPass::kUnpack:
      hipLaunchKernelGGL(k_mb_unpack, g, b, 0, s, boxes, first, nbox, nitems);
      break;
    case MBPass::kFx:
    case MBPass::kFz: {
      if (!mbox_supported(N)) return hipErrorInvalidValue;
      if (pass == MBPass::kFx && (maxM < 0 || maxM > kWTableMax || nb < 1 || nb > kFxMaxRadii))
        return hipErrorInvalidValue;
      if (pass == MBPass::kFz && (nb < 1 || nb > kFusedMaxRadii)) return hipErrorInvalidValue;
      mb_switch_host(N, [&](auto nc) {
        launch_mb_class<decltype(nc)::value>(pass, boxes, first, nbox, nitems, r0, nb, toff, maxM, g, rb, s);
      });
      break;
    }
    case MBPass::kSigma:
      hipLaunchKernelGGL(k_mb_sigma, dim3((unsigned)(nbox * kFusedMaxRadii * 8)), dim3(64), 0, s, boxes, nbox);
      break;
  }
  return hipGetLastError();
}

// ============================================================ launchers
#define PRC_FOR_EACH_N(X) X(16) X(32) X(64) X(128) X(192) X(256) X(512) X(1024) X(2048)

// per-stage twiddle table of plan P from FP64 sincospi, rounded to float
template <class P, int STAGE>
__device__ __forceinline__ bool twst_entry(int q, float2* out) {
  if constexpr (STAGE >= P::S) {
    return false;
  } else {
    constexpr int p = plan_p<P, STAGE>();
    constexpr int R = P::radices[STAGE];
    constexpr int off = plan_tw_offset<P, STAGE>();
    if (p > 1 && q >= off && q < off + p * R) {
      const int k = (q - off) / R, r = (q - off) % R;
      double sn, cs;
      sincospi(2.0 * (double)(r * k) / (double)(p * R), &sn, &cs);
      out[q] = make_float2((float)cs, (float)sn);
      return true;
    }
    return twst_entry<P, STAGE + 1>(q, out);
  }
}

template <class P>
__global__ void k_twst(float2* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < plan_tw_size<P>()) twst_entry<P, 0>(q, out);
}

template <int N>
static void twst_sizes(int* full, int* half) {
  *full = plan_tw_size<TilePlan<N>>();
  *half = plan_tw_size<Plan<N / 2>>();
}

template <int N>
static hipError_t twst_impl(float2* full, float2* half, hipStream_t s) {
  constexpr int a = plan_tw_size<TilePlan<N>>(), b = plan_tw_size<Plan<N / 2>>();
  if (a > 0) hipLaunchKernelGGL((k_twst<TilePlan<N>>), dim3((a + 255) / 256), dim3(256), 0, s, full);
  if (b > 0) hipLaunchKernelGGL((k_twst<Plan<N / 2>>), dim3((b + 255) / 256), dim3(256), 0, s, half);
  return hipGetLastError();
}

int stage_twiddle_size(int N, bool half) {
  int a = 0, b = 0;
  switch (N) {
#define X(NN)              \
  case NN:                 \
    twst_sizes<NN>(&a, &b); \
    break;
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return -1;
  }
  return half ? b : a;
}

int fx_twiddle_size(int N) {
  switch (N) {
#define X(NN) \
  case NN:    \
    return plan_tw_size<typename FxPlan<NN>::P>();
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return -1;
  }
}

template <int N>
static hipError_t fx_twst_impl(float2* out, hipStream_t s) {
  using P = typename FxPlan<N>::P;
  constexpr int a = plan_tw_size<P>();
  if (a > 0) hipLaunchKernelGGL((k_twst<P>), dim3((a + 255) / 256), dim3(256), 0, s, out);
  return hipGetLastError();
}

hipError_t launch_fx_twiddles(int N, float2* out, hipStream_t s) {
  switch (N) {
#define X(NN) \
  case NN:    \
    return fx_twst_impl<NN>(out, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_stage_twiddles(int N, float2* full, float2* half, hipStream_t s) {
  switch (N) {
#define X(NN) \
  case NN:    \
    return twst_impl<NN>(full, half, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

template <int N>
static hipError_t fwd_z_impl(const float* in, float2* out, float dinit, const float2* twh, const float2* twf,
                             hipStream_t s) {
  using P = Plan<N / 2>;
  constexpr int ROWS = 256 / P::T;
  const long long nrows = (long long)N * N;
  hipLaunchKernelGGL((k_fwd_z<P, Blk<N>::C>), dim3((unsigned)((nrows + ROWS - 1) / ROWS)), dim3(256), 0, s, in,
                     out, nrows, dinit, twh, twf);
  return hipGetLastError();
}

template <int N, class TS>
static hipError_t inv_z_impl(const TS* in, float* out, float iscale, const float2* twh, const float2* twf,
                             hipStream_t s) {
  using P = Plan<N / 2>;
  constexpr int ROWS = 256 / P::T;
  const long long nrows = (long long)N * N;
  hipLaunchKernelGGL((k_inv_z<P, Blk<N>::C, TS>), dim3((unsigned)((nrows + ROWS - 1) / ROWS)), dim3(256), 0, s, in,
                     out, nrows, iscale, twh, twf);
  return hipGetLastError();
}

template <int N, int DIR, class TI, class TO>
static hipError_t tile_impl(const TI* in, TO* out, const TileMap& mp, float oscale, const float2* tw,
                            hipStream_t s) {
  using P = TilePlan<N>;
  constexpr int C = Blk<N>::C;
  const long long ntiles = (long long)mp.nOuter * mp.nColBlocks;
  hipLaunchKernelGGL((k_tile<P, C, DIR, TI, TO>), dim3((unsigned)ntiles), dim3(P::T * C), 0, s, in, out, mp.in_sA,
                     mp.in_sO, mp.out_sA, mp.out_sO, mp.sCB, mp.nOuter, oscale, tw);
  return hipGetLastError();
}

// Blocks of kernel f (threads per block) resident on the current device at once
template <class F>
static int resident_blocks(F f, int threads) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, threads, 0);
  return (e == hipSuccess && cus > 0 && per_cu > 0) ? cus * per_cu : 1024;
}

template <int N, int DIR, class TS>
static hipError_t tile_ip_impl(TS* d, const TileMap& mp, float oscale, const float2* tw, hipStream_t s) {
  using P = TilePlan<N>;
  constexpr int C = Blk<N>::C;
  const long long ntiles = (long long)mp.nOuter * mp.nColBlocks;
  static const int resident = resident_blocks(k_tile_ip<P, C, DIR, TS>, P::T * C);
  const long long g = ntiles < resident ? ntiles : resident;
  hipLaunchKernelGGL((k_tile_ip<P, C, DIR, TS>), dim3((unsigned)g), dim3(P::T * C), 0, s, d, mp.in_sA, mp.in_sO,
                     mp.sCB, mp.nOuter, ntiles, oscale, tw);
  return hipGetLastError();
}

template <int N>
static hipError_t unpack_impl(const float2* S, float2* D0, float2* DN, hipStream_t s) {
  const long long n = (long long)N * N;
  hipLaunchKernelGGL((k_unpack<Blk<N>::C>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, S, D0, DN, N);
  return hipGetLastError();
}

template <int N>
static hipError_t disp_impl(const float2* S, const float2* D0, const float2* DN, float2* out, int axis, float dk,
                            float scale, hipStream_t s) {
  const long long n = (long long)N * N * (N / 2);
  long long g = (n + 255) / 256;
  if (g > (1 << 20)) g = 1 << 20;
  hipLaunchKernelGGL((k_disp<Blk<N>::C>), dim3((unsigned)g), dim3(256), 0, s, S, D0, DN, out, N, axis, dk, scale);
  return hipGetLastError();
}

template <int N>
static hipError_t pk_impl(const float2* S, const float2* D0, const float2* DN, float dk, float lnkmin, float dlnk,
                          int nk, double pscale, unsigned long long* count, double* sums, hipStream_t s) {
  const long long n = (long long)N * N * (N / 2 + 1);
  long long g = (n + 255) / 256;
  if (g > 2048) g = 2048;
  hipLaunchKernelGGL((k_pk<Blk<N>::C>), dim3((unsigned)g), dim3(256), 0, s, S, D0, DN, N, dk, lnkmin, dlnk, nk, pscale,
                     count, sums);
  return hipGetLastError();
}

template <int N>
static hipError_t filter_impl(const float2* S, const float2* D0, const float2* DN, float2* out, float dkR,
                              float invN3, hipStream_t s) {
  const long long n = (long long)N * N * (N / 2);
  long long g = (n + 255) / 256;
  if (g > (1 << 20)) g = 1 << 20;
  hipLaunchKernelGGL((k_filter<Blk<N>::C>), dim3((unsigned)g), dim3(256), 0, s, S, D0, DN, out, N, dkR, invN3);
  return hipGetLastError();
}

template <int N, class TS>
static hipError_t fx_impl(const float2* S, const float2* D0, const float2* DN, TS* Tbuf, float dk, float invN3,
                          const RadiusConst* rc, int r0, int nb, const float2* tw, const WTable& wt,
                          hipStream_t s) {
  using P = typename FxPlan<N>::P;
  constexpr int C = FxCols<N>::C, CF = FxCols<N>::CF;
  const long long nblk = (long long)N * ((N / 2) / C) * (C / CF);
  hipLaunchKernelGGL((k_fx<P, C, CF, TS>), dim3((unsigned)nblk), dim3(P::T * CF), sizeof(float4) * wt.M, s, S, D0,
                     DN, Tbuf, dk, invN3, rc, r0, nb, tw, wt.coef, wt.M, wt.invh);
  return hipGetLastError();
}

template <int N>
static int fz_grid(bool half, bool rb, bool fbar) {
  // grid-stride kernel: one block per resident slot of the device (all blocks
  // run in one wave of the grid), never more blocks than row groups
  using P = Plan<N / 2>;
  constexpr int C = Blk<N>::C;
  constexpr int ROWS = 256 / P::T;
  const long long need = ((long long)N * N + ROWS - 1) / ROWS;
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) {
    if (half)
      e = rb ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fz<P, C, true, __half2>, 256, 0)
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fz<P, C, false, __half2>, 256, 0);
    else if (fbar)
      e = rb ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fz<P, C, true, float2, true>, 256, 0)
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fz<P, C, false, float2, true>, 256, 0);
    else
      e = rb ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fz<P, C, true, float2>, 256, 0)
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fz<P, C, false, float2>, 256, 0);
  }
  long long g = (e == hipSuccess && cus > 0 && per_cu > 0) ? (long long)cus * per_cu : 2048;
  if (g > kFzMaxBlocks) g = kFzMaxBlocks;
  return (int)(need < g ? need : g);
}

template <int N, class TS>
static hipError_t fz_impl(const TS* Tbuf, float* zre, float* rbub, const RadiusConst* rc, int r0, int nb,
                          double* partial, int nblk, float iscale, const float2* twh, const float2* twf, bool fbar,
                          hipStream_t s) {
  using P = Plan<N / 2>;
  constexpr int C = Blk<N>::C;
  const long long nrows = (long long)N * N;
  if constexpr (std::is_same<TS, float2>::value) {
    if (fbar) {
      if (rbub)
        hipLaunchKernelGGL((k_fz<P, C, true, TS, true>), dim3(nblk), dim3(256), 0, s, Tbuf, zre, rbub, nrows, rc, r0,
                           nb, partial, iscale, twh, twf);
      else
        hipLaunchKernelGGL((k_fz<P, C, false, TS, true>), dim3(nblk), dim3(256), 0, s, Tbuf, zre, rbub, nrows, rc, r0,
                           nb, partial, iscale, twh, twf);
      return hipGetLastError();
    }
  }
  if (rbub)
    hipLaunchKernelGGL((k_fz<P, C, true, TS>), dim3(nblk), dim3(256), 0, s, Tbuf, zre, rbub, nrows, rc, r0, nb,
                       partial, iscale, twh, twf);
  else
    hipLaunchKernelGGL((k_fz<P, C, false, TS>), dim3(nblk), dim3(256), 0, s, Tbuf, zre, rbub, nrows, rc, r0, nb,
                       partial, iscale, twh, twf);
  return hipGetLastError();
}

int block_width(int N) {
  switch (N) {
#define X(NN) \
  case NN:    \
    return Blk<NN>::C;
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return 0;
  }
}

TileMap tile_map(int N, Axis axis) {
  const int C = block_width(N);
  const long long NC = (long long)N * C;
  TileMap m{};
  m.nOuter = N;
  m.nColBlocks = (N / 2) / C;
  m.sCB = (long long)N * N * C;
  switch (axis) {
    case Axis::kYInPlace:  // T layout (x, y): tile (x, kz block), transform along y
      m.in_sA = m.out_sA = C;
      m.in_sO = m.out_sO = NC;
      break;
    case Axis::kXTtoS:  // forward x: T layout (x, ky) -> S layout (ky, kx)
      m.in_sA = NC;
      m.in_sO = C;
      m.out_sA = C;
      m.out_sO = NC;
      break;
    case Axis::kXStoT:  // inverse x: S layout (ky, kx) -> T layout (x, ky)
      m.in_sA = C;
      m.in_sO = NC;
      m.out_sA = NC;
      m.out_sO = C;
      break;
  }
  return m;
}

hipError_t launch_fwd_z(int N, const float* in, float2* out, float dinit, const float2* twh, const float2* twf,
                        hipStream_t s) {
  switch (N) {
#define X(NN) \
  case NN:    \
    return fwd_z_impl<NN>(in, out, dinit, twh, twf, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_inv_z(int N, const float2* in, float* out, const float2* twh, const float2* twf, hipStream_t s) {
  switch (N) {
#define X(NN) \
  case NN:    \
    return inv_z_impl<NN, float2>(in, out, 1.0f, twh, twf, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_inv_z_h(int N, const void* in, float* out, float iscale, const float2* twh, const float2* twf,
                          hipStream_t s) {
  switch (N) {
#define X(NN) \
  case NN:    \
    return inv_z_impl<NN, __half2>(static_cast<const __half2*>(in), out, iscale, twh, twf, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_tile(int N, int dir, const float2* in, float2* out, const TileMap& mp, const float2* tw,
                       hipStream_t s) {
  // The in-place inverse y pass at N = 1024 and 2048 runs the persistent
  // kernel (the next tile's 16 FP32 pairs per thread loaded into registers
  // while the current one is transformed: 112-113 VGPRs, no spills): same
  // bits, 141.1-141.5 -> 134.4-138.2 ms per 1024^3 x 96 step and 1548-1555
  // -> 1457-1459 ms per 2048^3 x 128 step.  Other sizes keep one tile per
  // block (the persistent kernel measured slower for FP32 pencils of 8 values
  // at N = 512).
  if ((N == 1024 || N == 2048) && dir > 0 && in == out && mp.in_sA == mp.out_sA && mp.in_sO == mp.out_sO)
    return N == 1024 ? tile_ip_impl<1024, +1, float2>(out, mp, 1.0f, tw, s)
                     : tile_ip_impl<2048, +1, float2>(out, mp, 1.0f, tw, s);
  switch (N) {
#define X(NN)                                                                          \
  case NN:                                                                             \
    return dir > 0 ? tile_impl<NN, +1, float2, float2>(in, out, mp, 1.0f, tw, s)      \
                   : tile_impl<NN, -1, float2, float2>(in, out, mp, 1.0f, tw, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_tile_h(int N, void* data, const TileMap& mp, float oscale, const float2* tw, hipStream_t s) {
  __half2* d = static_cast<__half2*>(data);
  // in place along y only (the persistent kernel reads and writes one map)
  if (mp.in_sA != mp.out_sA || mp.in_sO != mp.out_sO) return hipErrorInvalidValue;
  switch (N) {
#define X(NN) \
  case NN:    \
    return tile_ip_impl<NN, +1, __half2>(d, mp, oscale, tw, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_unpack(int N, const float2* S, float2* D0, float2* DN, hipStream_t s) {
  switch (N) {
#define X(NN) \
  case NN:    \
    return unpack_impl<NN>(S, D0, DN, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_filter(int N, const float2* S, const float2* D0, const float2* DN, float2* out, float dkR,
                         float invN3, hipStream_t s) {
  switch (N) {
#define X(NN) \
  case NN:    \
    return filter_impl<NN>(S, D0, DN, out, dkR, invN3, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_pk(int N, const float2* S, const float2* D0, const float2* DN, float dk, float lnkmin, float dlnk,
                     int nk, double pscale, unsigned long long* count, double* sums, hipStream_t s) {
  if (nk < 1 || nk > kPkMaxBins) return hipErrorInvalidValue;
  hipError_t e = hipMemsetAsync(count, 0, sizeof(unsigned long long) * nk, s);
  if (e == hipSuccess) e = hipMemsetAsync(sums, 0, sizeof(double) * 3 * nk, s);
  if (e != hipSuccess) return e;
  switch (N) {
#define X(NN) \
  case NN:    \
    return pk_impl<NN>(S, D0, DN, dk, lnkmin, dlnk, nk, pscale, count, sums, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_disp(int N, const float2* S, const float2* D0, const float2* DN, float2* out, int axis, float dk,
                       float scale, hipStream_t s) {
  if (axis < 0 || axis > 2) return hipErrorInvalidValue;
  switch (N) {
#define X(NN) \
  case NN:    \
    return disp_impl<NN>(S, D0, DN, out, axis, dk, scale, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_fx(int N, const float2* S, const float2* D0, const float2* DN, void* Tbuf, bool half, float dk,
                     float invN3, const RadiusConst* rc, int r0, int nb, const float2* tw, const WTable& wt,
                     hipStream_t s) {
  if (nb < 1 || nb > kFxMaxRadii || wt.M < 0 || wt.M > kWTableMax) return hipErrorInvalidValue;
  switch (N) {
#define X(NN)                                                                                                  \
  case NN:                                                                                                     \
    return half ? fx_impl<NN, __half2>(S, D0, DN, static_cast<__half2*>(Tbuf), dk, invN3, rc, r0, nb, tw, wt, s) \
                : fx_impl<NN, float2>(S, D0, DN, static_cast<float2*>(Tbuf), dk, invN3, rc, r0, nb, tw, wt, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

int fz_blocks(int N, bool half, bool rb, bool fbar) {
  switch (N) {
#define X(NN) \
  case NN:    \
    return fz_grid<NN>(half, rb, fbar);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return 0;
  }
}

hipError_t launch_fz(int N, const void* Tbuf, bool half, float* zre, float* rbub, const RadiusConst* rc, int r0,
                     int nb, double* partial, int nblk, float iscale, const float2* twh, const float2* twf, bool fbar,
                     hipStream_t s) {
  if (nb < 1 || nb > kFusedMaxRadii) return hipErrorInvalidValue;
  switch (N) {
#define X(NN)                                                                                                 \
  case NN:                                                                                                    \
    return half ? fz_impl<NN, __half2>(static_cast<const __half2*>(Tbuf), zre, rbub, rc, r0, nb, partial, nblk, \
                                       iscale, twh, twf, false, s)                                            \
                : fz_impl<NN, float2>(static_cast<const float2*>(Tbuf), zre, rbub, rc, r0, nb, partial, nblk,   \
                                      iscale, twh, twf, fbar, s);
    PRC_FOR_EACH_N(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace prc
