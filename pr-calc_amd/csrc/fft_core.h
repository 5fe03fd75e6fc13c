// Device building blocks for the pencil FFT passes (gfx950, wave64).
//
// A pencil of length L is owned by T = L/E threads; thread i holds the E
// elements x[i + e*T], e = 0..E-1, in registers.  Each pass runs a
// self-sorting (Stockham) sequence of radix stages; between stages the
// pencil is exchanged through LDS.  With this register layout the first
// stage reads x[i + e*T] and the last stage produces y[i + e*T] in natural
// order, so global loads and stores are coalesced along whichever memory axis
// the caller maps consecutive lanes to.
//
// Stage (p, R), p = product of the radices already applied:
//   butterfly b = i + m*T (m < E/R) takes x[b + r*L/R] = v[m + r*E/R],
//   multiplies by w_{pR}^{r*(b mod p)}, applies the R-point DFT and writes
//   y[(b/p)*p*R + (b mod p) + r*p].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace prc {

// LDS loads whose values are first used after a block barrier must still be
// done before it when other waves write the same LDS after that barrier.  The
// backend may sink such loads below __syncthreads (seen in the persistent y
// pass: the last gather of one tile moved past the barrier that frees the
// buffer, so that other waves' writes of the next tile raced with it, and a
// few waves of one tile came out wrong in about 5% of runs at 128^3).
// lds_hold(v) before the barrier takes the values as inputs of an empty
// volatile asm, which keeps their loads (and what they feed) above it.
template <int E>
__device__ __forceinline__ void lds_hold(const float2 (&v)[E]) {
#pragma unroll
  for (int e = 0; e < E; ++e) asm volatile("" ::"v"(v[e].x), "v"(v[e].y));
}

// floor(log2(n)) for n >= 1
constexpr int ilog2(int n) { return n > 1 ? 1 + ilog2(n / 2) : 0; }

// ---------------------------------------------------------------- complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// a * (DIR * i)
template <int DIR>
__device__ __forceinline__ float2 mul_di(float2 a) {
  return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ------------------------------------------------- compile-time trigonometry
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int n = 1; n < 14; ++n) { term *= -x2 / ((2 * n) * (2 * n + 1)); sum += term; }
  return sum;
}
constexpr double taylor_cos(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int n = 1; n < 14; ++n) { term *= -x2 / ((2 * n - 1) * (2 * n)); sum += term; }
  return sum;
}
// cos(2 pi m / R) and sin(2 pi m / R) for integers, reduced by octant so the
// series argument stays within pi/4.
constexpr void twiddle_cs(long m, long R, double& c, double& s) {
  m %= R;
  if (m < 0) m += R;
  // angle = 2 pi m / R; octant o = floor(8 m / R)
  long o = (8 * m) / R;
  // residual angle measured from the octant's reference axis
  double frac = (8.0 * (double)m - (double)(o * R)) / (double)R;  // in [0,1)
  double a = frac * (kPi / 4);                                    // in [0, pi/4)
  double ca = taylor_cos(a), sa = taylor_sin(a);
  double cb = taylor_cos(kPi / 4 - a), sb = taylor_sin(kPi / 4 - a);
  switch (o) {
    case 0: c = ca; s = sa; break;
    case 1: c = sb; s = cb; break;
    case 2: c = -sa; s = ca; break;
    case 3: c = -cb; s = sb; break;
    case 4: c = -ca; s = -sa; break;
    case 5: c = -sb; s = -cb; break;
    case 6: c = sa; s = -ca; break;
    default: c = cb; s = -sb; break;
  }
}
constexpr double tw_c(long m, long R) { double c = 0, s = 0; twiddle_cs(m, R, c, s); return c; }
constexpr double tw_s(long m, long R) { double c = 0, s = 0; twiddle_cs(m, R, c, s); return s; }

// v * exp(DIR * 2 pi i m / R) with m, R compile-time constants
template <int DIR, long M, long R>
__device__ __forceinline__ float2 twc(float2 v) {
  constexpr long m = ((M % R) + R) % R;
  if constexpr (m == 0) {
    return v;
  } else if constexpr (4 * m == R) {
    return mul_di<DIR>(v);
  } else if constexpr (2 * m == R) {
    return make_float2(-v.x, -v.y);
  } else if constexpr (4 * m == 3 * R) {
    return mul_di<-DIR>(v);
  } else {
    constexpr float c = (float)tw_c(m, R);
    constexpr float s = (float)(DIR * tw_s(m, R));
    return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
  }
}

// ------------------------------------------------ in-register small DFTs
// y[k] = sum_n x[n] exp(DIR 2 pi i n k / R), in place, natural order.
template <int R, int DIR>
struct DFT;

template <int DIR>
struct DFT<1, DIR> {
  __device__ __forceinline__ static void run(float2*) {}
};

template <int DIR>
struct DFT<2, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int DIR>
struct DFT<3, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    constexpr float kS = 0.86602540378443864676f;  // sqrt(3)/2
    float2 s = cadd(v[1], v[2]);
    float2 d = csub(v[1], v[2]);
    float2 y0 = cadd(v[0], s);
    float2 t1 = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
    float2 t2 = cscale(mul_di<DIR>(d), kS);
    v[0] = y0;
    v[1] = cadd(t1, t2);
    v[2] = csub(t1, t2);
  }
};

template <int DIR>
struct DFT<4, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 t0 = cadd(v[0], v[2]);
    float2 t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]);
    float2 t3 = mul_di<DIR>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

constexpr int first_factor(int R) { return (R % 4 == 0 && R != 4) ? 4 : (R % 2 == 0 ? 2 : 3); }

// Composite R = R1 * R2: n = n1 + R1 n2, k = k2 + R2 k1.
template <int R, int DIR>
struct DFT {
  static constexpr int R1 = first_factor(R);
  static constexpr int R2 = R / R1;
  template <int N1, int K2>
  __device__ __forceinline__ static float2 tw(float2 v) { return twc<DIR, (long)N1 * K2, R>(v); }

  template <int N1>
  __device__ __forceinline__ static void inner(const float2* v, float2 (&a)[R1][R2]) {
    float2 t[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) t[n2] = v[N1 + R1 * n2];
    DFT<R2, DIR>::run(t);
    twiddle_row<N1, 0>(t, a);
  }
  template <int N1, int K2>
  __device__ __forceinline__ static void twiddle_row(const float2* t, float2 (&a)[R1][R2]) {
    if constexpr (K2 < R2) {
      a[N1][K2] = tw<N1, K2>(t[K2]);
      twiddle_row<N1, K2 + 1>(t, a);
    }
  }
  template <int N1>
  __device__ __forceinline__ static void all_inner(const float2* v, float2 (&a)[R1][R2]) {
    if constexpr (N1 < R1) {
      inner<N1>(v, a);
      all_inner<N1 + 1>(v, a);
    }
  }
  __device__ __forceinline__ static void run(float2* v) {
    float2 a[R1][R2];
    all_inner<0>(v, a);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      float2 t[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) t[n1] = a[n1][k2];
      DFT<R1, DIR>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) v[k2 + R2 * k1] = t[k1];
    }
  }
};

// --------------------------------------------------------------- plans
// A plan: transform length L, E elements per thread (T = L/E threads per
// pencil) and the radix sequence; every radix divides E.
template <int L_, int E_, int... R>
struct PlanT {
  static constexpr int L = L_, E = E_, T = L_ / E_;
  static constexpr int radices[] = {R...};
  static constexpr int S = sizeof...(R);
};

// Default plans (one 64-lane wave or less per pencil for the z passes).
template <int L>
struct Plan;
template <> struct Plan<8> : PlanT<8, 8, 8> {};
template <> struct Plan<16> : PlanT<16, 16, 16> {};
template <> struct Plan<32> : PlanT<32, 8, 8, 4> {};
template <> struct Plan<64> : PlanT<64, 8, 8, 8> {};
template <> struct Plan<96> : PlanT<96, 24, 3, 8, 4> {};
template <> struct Plan<128> : PlanT<128, 16, 16, 8> {};
template <> struct Plan<192> : PlanT<192, 24, 3, 8, 8> {};
template <> struct Plan<256> : PlanT<256, 16, 16, 16> {};
template <> struct Plan<512> : PlanT<512, 8, 8, 8, 8> {};
template <> struct Plan<1024> : PlanT<1024, 16, 16, 16, 4> {};
template <> struct Plan<2048> : PlanT<2048, 16, 16, 16, 8> {};

// Plans of the tiled column passes (forward and inverse x, y; the fused x
// pass): 16 threads per pencil up to N = 256, so that a tile of 16 columns is
// one 256-thread block for every N of that range (the multi-box kernels run
// boxes of several sizes in one launch of 256-thread blocks).
template <int L>
struct TilePlan : Plan<L> {};
template <> struct TilePlan<64> : PlanT<64, 4, 4, 4, 4> {};
template <> struct TilePlan<128> : PlanT<128, 8, 8, 4, 4> {};
template <> struct TilePlan<192> : PlanT<192, 12, 3, 4, 4, 4> {};

// Wider plans for kernels that keep a second register copy of their tile
// (the fused filter pass): fewer threads per pencil, more registers each.
template <int L>
struct PlanWide : Plan<L> {};
template <> struct PlanWide<1024> : PlanT<1024, 32, 32, 32> {};
template <> struct PlanWide<2048> : PlanT<2048, 32, 32, 16, 4> {};

template <class P, int S>
constexpr int plan_p() {  // product of the first S radices
  int p = 1;
  for (int s = 0; s < S; ++s) p *= P::radices[s];
  return p;
}

// Per-stage twiddle tables: stage s with p > 1 owns p*R entries
// w[k*R + r] = exp(+2 pi i r k / (p R)), k < p, r < R, stored back to back in
// stage order, so one butterfly reads its R twiddles from consecutive addresses.
template <class P, int S>
constexpr int plan_tw_offset() {
  int off = 0, p = 1;
  for (int s = 0; s < S; ++s) {
    if (p > 1) off += p * P::radices[s];
    p *= P::radices[s];
  }
  return off;
}
template <class P>
constexpr int plan_tw_size() {
  return plan_tw_offset<P, P::S>();
}

// Synchronisation between the LDS phases of a transform.  Pencils spread over
// the waves of a block need a block barrier.  A pencil held by the lanes of
// one wave only needs the compiler to keep its LDS accesses in program order:
// the LDS performs the operations of one wave in order, so a wave-scope fence
// (no instruction) plus a scheduling barrier suffices.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS view of one pencil: element a at base[a * sa]
struct LdsPencil {
  static constexpr bool kWaveLocal = false;
  float2* base;
  int sa;
  __device__ __forceinline__ float2& operator[](int a) const { return base[a * sa]; }
};

// LDS layout of pencils that each own a contiguous row (the z passes): row rr
// starts at rr * (L + PAD) and element a sits at a ^ ((a >> SH) & MK), an XOR
// swizzle inside aligned 32-element blocks.  The early Stockham stages write
// with a stride of R elements, which without it lands all lanes of a wave on a
// few banks; SH/MK/PAD are chosen per length so that every write and gather of
// every stage touches distinct banks for 16 (32-bank) or 32 (64-bank) lanes of
// float2 per cycle.  L must be a multiple of 32 when MK > 15.
template <int L>
struct RowSwizzle { static constexpr int SH = 0, MK = 0, PAD = 0; };
template <> struct RowSwizzle<32> { static constexpr int SH = 3, MK = 15, PAD = 4; };
template <> struct RowSwizzle<64> { static constexpr int SH = 3, MK = 15, PAD = 8; };
template <> struct RowSwizzle<96> { static constexpr int SH = 0, MK = 0, PAD = 4; };
template <> struct RowSwizzle<128> { static constexpr int SH = 4, MK = 15, PAD = 8; };
template <> struct RowSwizzle<192> { static constexpr int SH = 0, MK = 0, PAD = 8; };
template <> struct RowSwizzle<256> { static constexpr int SH = 4, MK = 15, PAD = 16; };
template <> struct RowSwizzle<512> { static constexpr int SH = 3, MK = 31, PAD = 0; };
template <> struct RowSwizzle<1024> { static constexpr int SH = 4, MK = 31, PAD = 0; };

// LDS view of one column of a tile whose columns are interleaved (element a
// of column c at a * CF + c): rows are padded by one row of CF every 2^SH rows,
// so that two pencils 2^SH rows apart (the stride of a radix-2^SH first stage)
// sit half a bank cycle apart.  Holds (L + L / 2^SH) * CF elements.
template <int SH>
struct LdsColPad {
  static constexpr bool kWaveLocal = false;
  float2* base;  // column start
  int sa;        // CF
  __device__ __forceinline__ float2& operator[](int a) const { return base[(a + (a >> SH)) * sa]; }
};

// With WAVE the pencils must be held by the lanes of one wave (T <= 64
// consecutive lanes, 64 % T == 0) and the transforms synchronise per wave
// instead of per block.
template <int L, bool WAVE = false>
struct LdsRow {
  static constexpr bool kWaveLocal = WAVE;
  using Sw = RowSwizzle<L>;
  static constexpr int STRIDE = L + Sw::PAD;  // float2 per row
  float2* base;
  __device__ __forceinline__ float2& operator[](int a) const {
    if constexpr (Sw::MK == 0) {
      return base[a];
    } else {
      return base[a ^ ((a >> Sw::SH) & Sw::MK)];
    }
  }
};

// One Stockham stage on registers v (thread index i within the pencil); tw is
// the plan's per-stage twiddle table.
// If LAST, outputs stay in registers (natural order, v[e] = y[i + e*T]);
// otherwise they are written to LDS at their Stockham positions.
template <class P, int STAGE, int DIR, bool LAST, class V, bool TWT = false>
__device__ __forceinline__ void stockham_stage(float2 (&v)[P::E], int i, const float2* __restrict__ tw,
                                               V lds) {
  constexpr int L = P::L, E = P::E, T = P::T;
  constexpr int R = P::radices[STAGE];
  constexpr int p = plan_p<P, STAGE>();
  constexpr int NB = E / R;  // butterflies per thread
  constexpr int TWO = plan_tw_offset<P, STAGE>();
  static_assert(E % R == 0, "radix must divide E");
  (void)L;
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    float2 u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[m + r * NB];
    const int b = i + m * T;
    const int k = (p == 1) ? 0 : (b % p);
    if constexpr (p > 1) {
      // TWT: the stage's table is stored transposed, entry (k, r) at r*p + k
      const float2* wk = TWT ? tw + TWO + k : tw + TWO + k * R;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = TWT ? wk[r * p] : wk[r];
        if (DIR < 0) w.y = -w.y;
        u[r] = cmul(u[r], w);
      }
    }
    DFT<R, DIR>::run(u);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[m + r * NB] = u[r];
    } else {
      const int j = (b - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[j + r * p] = u[r];
    }
  }
}

template <class P, class V>
__device__ __forceinline__ void lds_gather(float2 (&v)[P::E], int i, V lds) {
#pragma unroll
  for (int e = 0; e < P::E; ++e) v[e] = lds[i + e * P::T];
}

// Full pencil transform: registers in natural layout -> registers in natural
// layout.  All threads of the block (of the wave, for a wave-local view) must
// call it (it contains barriers); the LDS region may be reused right after it
// returns.  TAIL = false leaves out the barrier after the last gather: the
// caller then synchronises before anything rewrites the region (a loop over
// tiles or radii syncs at the top of its body, which every wave reaches only
// after its last stage has used the gathered values).
template <class P, int DIR, int STAGE = 0, class V, bool TWT = false, bool TAIL = true>
__device__ __forceinline__ void pencil_fft(float2 (&v)[P::E], int i, const float2* __restrict__ tw, V lds) {
  if constexpr (STAGE == P::S - 1) {
    stockham_stage<P, STAGE, DIR, true, V, TWT>(v, i, tw, lds);
  } else {
    static_assert(!V::kWaveLocal || (P::T <= 64 && 64 % P::T == 0), "wave-local pencil must fit in one wave");
    stockham_stage<P, STAGE, DIR, false, V, TWT>(v, i, tw, lds);
    if constexpr (V::kWaveLocal) wave_lds_sync(); else __syncthreads();
    lds_gather<P>(v, i, lds);
    if constexpr (TAIL || STAGE + 2 < P::S) {
      if constexpr (!V::kWaveLocal) lds_hold(v);  // the gather is done before the barrier
      if constexpr (V::kWaveLocal) wave_lds_sync(); else __syncthreads();
    }
    pencil_fft<P, DIR, STAGE + 1, V, TWT, TAIL>(v, i, tw, lds);
  }
}

// Copy plan P's per-stage twiddle table src ([k][r] per stage) to dst with
// every stage transposed ([r][k]), threads t = 0..nt-1 of the block.
template <class P, int STAGE = 0>
__device__ __forceinline__ void copy_twiddles_transposed(float2* dst, const float2* __restrict__ src, int t,
                                                         int nt) {
  if constexpr (STAGE < P::S) {
    constexpr int p = plan_p<P, STAGE>();
    constexpr int R = P::radices[STAGE];
    constexpr int TWO = plan_tw_offset<P, STAGE>();
    if constexpr (p > 1) {
      for (int q = t; q < p * R; q += nt) {
        const int k = q / R, r = q % R;
        dst[TWO + r * p + k] = src[TWO + q];
      }
    }
    copy_twiddles_transposed<P, STAGE + 1>(dst, src, t, nt);
  }
}

// Mirror element x[(L - k) mod L] of every register x[k], k = i + e*T, for
// pencils held by T <= 64 consecutive lanes of a wave: for i > 0 it sits in
// lane (T - i) of the pencil at register E-1-e, for i = 0 in this lane at
// register (E - e) mod E.  Cross-lane moves use ds_bpermute (no LDS storage).
template <class P>
__device__ __forceinline__ void mirror_lanes(const float2 (&v)[P::E], float2 (&m)[P::E], int i) {
  constexpr int E = P::E, T = P::T;
  static_assert(T <= 64 && 64 % T == 0, "pencil must fit in one wave");
  const int lane = __lane_id();
  const int src = (lane - i) + ((T - i) % T);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float2 w = v[E - 1 - e];
    float2 got;
    got.x = __shfl(w.x, src, 64);
    got.y = __shfl(w.y, src, 64);
    m[e] = (i == 0) ? v[(E - e) % E] : got;
  }
}

// The same for 32-bit words (FP16 complex pairs): one move per register
template <class P>
__device__ __forceinline__ void mirror_lanes(const uint32_t (&v)[P::E], uint32_t (&m)[P::E], int i) {
  constexpr int E = P::E, T = P::T;
  static_assert(T <= 64 && 64 % T == 0, "pencil must fit in one wave");
  const int lane = __lane_id();
  const int src = (lane - i) + ((T - i) % T);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const uint32_t got = (uint32_t)__shfl((int)v[E - 1 - e], src, 64);
    m[e] = (i == 0) ? v[(E - e) % E] : got;
  }
}

}  // namespace prc
