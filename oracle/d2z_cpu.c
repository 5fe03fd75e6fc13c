/* d2z_cpu.c -- C restatement of pr-calc's delta2zreion loop in the
 * reference's own FP32 arithmetic (marcelo-alvarez/pr-calc).
 *
 * TEST INFRASTRUCTURE ONLY: the CPU baseline of bench.py and a second checker
 * for the tests.  The product never links or runs it.
 *
 * Follows, step by step:
 *   main_d2z.C:57-59        delta /= DInit over the padded array (float)
 *   main_d2z.C:67           unnormalised in-place r2c (FFTW layout [N][N][N+2])
 *   delta2zreion_d2z.C:58-89  float k grid, W = 3.*(sin(rkR)-rkR*cos(rkR))/pow(rkR,3)
 *                           with float sin/cos and float difference, k=0 -> 0
 *   delta2zreion_d2z.C:92   c2r
 *   arrayoperations.C:63-68 NormalizeArray: array[i] /= pow((float)N,3)
 *   arrayoperations.C:4-35  GetMeanSigma in double
 *   delta2zreion_d2z.C:106-146,206-216  z = A*(d + sqrt(2.*(smin^2-s^2))*erfizeta) - 1
 *                           in double, stored float, strict running maximum
 * The FFTW3 transforms are replaced by this file's own FP32 Stockham FFT
 * (radix 2/3/4) -- FFTW is not available here; its rounding pattern is not
 * reproduced.  OpenMP threads over pencils stand in for the MPI slabs.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
  float re, im;
} cpx;

/* ------------------------------------------------------------------ 1D FFT */
typedef struct {
  int n;
  int nst;
  int rad[32];
  cpx* tw; /* tw[m] = exp(+2 pi i m / n) */
} plan1d;

static int plan_init(plan1d* p, int n) {
  p->n = n;
  p->nst = 0;
  int m = n;
  while (m % 4 == 0) { p->rad[p->nst++] = 4; m /= 4; }
  while (m % 2 == 0) { p->rad[p->nst++] = 2; m /= 2; }
  while (m % 3 == 0) { p->rad[p->nst++] = 3; m /= 3; }
  if (m != 1) return -1;
  p->tw = (cpx*)malloc(sizeof(cpx) * n);
  for (int k = 0; k < n; ++k) {
    double a = 2.0 * M_PI * (double)k / (double)n;
    p->tw[k].re = (float)cos(a);
    p->tw[k].im = (float)sin(a);
  }
  return 0;
}

static void plan_free(plan1d* p) { free(p->tw); }

/* One transform of length n, sign = +1 (inverse) or -1 (forward); x is
 * overwritten with the result, y is scratch of length n. */
static void fft1d(const plan1d* pl, int sign, cpx* x, cpx* y) {
  const int n = pl->n;
  int p = 1;
  cpx* src = x;
  cpx* dst = y;
  for (int s = 0; s < pl->nst; ++s) {
    const int R = pl->rad[s];
    const int T = n / R;
    const int tws = n / (p * R);
    for (int i = 0; i < T; ++i) {
      const int k = i % p;
      const int j = (i - k) * R + k;
      cpx u[4] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
      for (int r = 0; r < R; ++r) {
        cpx v = src[i + r * T];
        if (r && k) {
          cpx w = pl->tw[r * k * tws];
          float wi = sign > 0 ? w.im : -w.im;
          float re = v.re * w.re - v.im * wi;
          float im = v.re * wi + v.im * w.re;
          v.re = re;
          v.im = im;
        }
        u[r] = v;
      }
      if (R == 2) {
        cpx a = u[0], b = u[1];
        u[0].re = a.re + b.re; u[0].im = a.im + b.im;
        u[1].re = a.re - b.re; u[1].im = a.im - b.im;
      } else if (R == 4) {
        cpx t0 = {u[0].re + u[2].re, u[0].im + u[2].im};
        cpx t1 = {u[0].re - u[2].re, u[0].im - u[2].im};
        cpx t2 = {u[1].re + u[3].re, u[1].im + u[3].im};
        cpx d = {u[1].re - u[3].re, u[1].im - u[3].im};
        cpx t3 = sign > 0 ? (cpx){-d.im, d.re} : (cpx){d.im, -d.re};
        u[0].re = t0.re + t2.re; u[0].im = t0.im + t2.im;
        u[2].re = t0.re - t2.re; u[2].im = t0.im - t2.im;
        u[1].re = t1.re + t3.re; u[1].im = t1.im + t3.im;
        u[3].re = t1.re - t3.re; u[3].im = t1.im - t3.im;
      } else { /* R == 3 */
        const float c = 0.86602540378443864676f;
        cpx s = {u[1].re + u[2].re, u[1].im + u[2].im};
        cpx d = {u[1].re - u[2].re, u[1].im - u[2].im};
        cpx y0 = {u[0].re + s.re, u[0].im + s.im};
        cpx t1 = {u[0].re - 0.5f * s.re, u[0].im - 0.5f * s.im};
        cpx t2 = sign > 0 ? (cpx){-d.im * c, d.re * c} : (cpx){d.im * c, -d.re * c};
        u[0] = y0;
        u[1].re = t1.re + t2.re; u[1].im = t1.im + t2.im;
        u[2].re = t1.re - t2.re; u[2].im = t1.im - t2.im;
      }
      for (int r = 0; r < R; ++r) dst[j + r * p] = u[r];
    }
    p *= R;
    cpx* t = src;
    src = dst;
    dst = t;
  }
  if (src != x) memcpy(x, src, sizeof(cpx) * n);
}

/* ----------------------------------------------------------------- handle */
typedef struct {
  int N;
  float box;         /* BoxSize/h, Mpc */
  long nreal;        /* N*N*(N+2) */
  float* work;       /* r2c in place: delta, then its half spectrum */
  float* smooth;     /* delta_R */
  float* zreion;     /* padded, zero-initialised (allocate_d2z.C:37) */
  plan1d pN, pH;     /* length N (x, y), length N/2 (z real transform) */
  cpx* twN;          /* exp(+2 pi i k/N) for the real-transform untangling */
  int nthreads;
} d2z_cpu;

void* d2z_cpu_create(int N, float box_mpc, int nthreads) {
  d2z_cpu* h = (d2z_cpu*)calloc(1, sizeof(d2z_cpu));
  h->N = N;
  h->box = box_mpc;
  h->nreal = (long)N * N * (N + 2);
  h->work = (float*)malloc(sizeof(float) * h->nreal);
  h->smooth = (float*)malloc(sizeof(float) * h->nreal);
  h->zreion = (float*)calloc(h->nreal, sizeof(float));
  h->nthreads = nthreads > 0 ? nthreads : 1;
  if (!h->work || !h->smooth || !h->zreion || plan_init(&h->pN, N) || plan_init(&h->pH, N / 2)) {
    free(h->work); free(h->smooth); free(h->zreion); free(h);
    return NULL;
  }
  h->twN = h->pN.tw;
  return h;
}

void d2z_cpu_destroy(void* hv) {
  d2z_cpu* h = (d2z_cpu*)hv;
  if (!h) return;
  free(h->work); free(h->smooth); free(h->zreion);
  plan_free(&h->pN); plan_free(&h->pH);
  free(h);
}

/* z-axis real transforms on each padded row (N reals <-> N/2+1 complex) */
static void rows_r2c(d2z_cpu* h, float* a) {
  const int N = h->N, L = N / 2;
#pragma omp parallel num_threads(h->nthreads)
  {
    cpx* z = (cpx*)malloc(sizeof(cpx) * L);
    cpx* s = (cpx*)malloc(sizeof(cpx) * L);
#pragma omp for schedule(static)
    for (long row = 0; row < (long)N * N; ++row) {
      float* r = a + row * (N + 2);
      for (int n = 0; n < L; ++n) { z[n].re = r[2 * n]; z[n].im = r[2 * n + 1]; }
      fft1d(&h->pH, -1, z, s);
      cpx* X = (cpx*)r;
      X[0].re = z[0].re + z[0].im; X[0].im = 0.f;
      X[L].re = z[0].re - z[0].im; X[L].im = 0.f;
      for (int k = 1; k < L; ++k) {
        cpx Z = z[k], Zm = {z[L - k].re, -z[L - k].im};
        cpx E = {0.5f * (Z.re + Zm.re), 0.5f * (Z.im + Zm.im)};
        cpx D = {0.5f * (Z.re - Zm.re), 0.5f * (Z.im - Zm.im)};
        cpx O = {D.im, -D.re};
        cpx w = {h->twN[k].re, -h->twN[k].im};
        X[k].re = E.re + (w.re * O.re - w.im * O.im);
        X[k].im = E.im + (w.re * O.im + w.im * O.re);
      }
    }
    free(z); free(s);
  }
}

static void rows_c2r(d2z_cpu* h, float* a) {
  const int N = h->N, L = N / 2;
#pragma omp parallel num_threads(h->nthreads)
  {
    cpx* z = (cpx*)malloc(sizeof(cpx) * L);
    cpx* s = (cpx*)malloc(sizeof(cpx) * L);
#pragma omp for schedule(static)
    for (long row = 0; row < (long)N * N; ++row) {
      float* r = a + row * (N + 2);
      const cpx* X = (const cpx*)r;
      for (int k = 0; k < L; ++k) {
        cpx a0 = X[k], b = {X[L - k].re, -X[L - k].im};
        cpx sm = {a0.re + b.re, a0.im + b.im}, d = {a0.re - b.re, a0.im - b.im};
        cpx w = h->twN[k];
        cpx wd = {w.re * d.re - w.im * d.im, w.re * d.im + w.im * d.re};
        z[k].re = sm.re - wd.im;
        z[k].im = sm.im + wd.re;
      }
      fft1d(&h->pH, +1, z, s);
      for (int n = 0; n < L; ++n) { r[2 * n] = z[n].re; r[2 * n + 1] = z[n].im; }
    }
    free(z); free(s);
  }
}

/* complex transform along x (axis 0) or y (axis 1) of the [N][N][N/2+1] view */
static void axis_fft(d2z_cpu* h, float* a, int axis, int sign) {
  const int N = h->N, Nc = N / 2 + 1;
  const long sA = axis == 0 ? (long)N * Nc : Nc; /* complex stride along the axis */
  const long sO = axis == 0 ? Nc : (long)N * Nc; /* stride of the other outer index */
  const int CB = 8;
  const int ncb = (Nc + CB - 1) / CB;
  cpx* c = (cpx*)a;
#pragma omp parallel num_threads(h->nthreads)
  {
    cpx* buf = (cpx*)malloc(sizeof(cpx) * N * CB);
    cpx* s = (cpx*)malloc(sizeof(cpx) * N);
#pragma omp for schedule(static)
    for (long t = 0; t < (long)N * ncb; ++t) {
      const long o = t / ncb;
      const int c0 = (int)(t % ncb) * CB;
      const int nc = (c0 + CB <= Nc) ? CB : Nc - c0;
      cpx* base = c + o * sO + c0;
      for (int i = 0; i < N; ++i)
        for (int q = 0; q < nc; ++q) buf[q * N + i] = base[i * sA + q];
      for (int q = 0; q < nc; ++q) fft1d(&h->pN, sign, buf + q * N, s);
      for (int i = 0; i < N; ++i)
        for (int q = 0; q < nc; ++q) base[i * sA + q] = buf[q * N + i];
    }
    free(buf); free(s);
  }
}

static double now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* delta: dense N^3; returns seconds */
double d2z_cpu_forward(void* hv, const float* delta, float DInit) {
  d2z_cpu* h = (d2z_cpu*)hv;
  const int N = h->N;
  double t0 = now();
#pragma omp parallel for num_threads(h->nthreads) schedule(static)
  for (long row = 0; row < (long)N * N; ++row) {
    float* r = h->work + row * (N + 2);
    memcpy(r, delta + row * N, sizeof(float) * N);
    r[N] = r[N + 1] = 0.f;
    for (int k = 0; k < N + 2; ++k) r[k] /= DInit;
  }
  rows_r2c(h, h->work);
  axis_fft(h, h->work, 1, -1);
  axis_fft(h, h->work, 0, -1);
  return now() - t0;
}

void d2z_cpu_reset(void* hv) {
  d2z_cpu* h = (d2z_cpu*)hv;
  memset(h->zreion, 0, sizeof(float) * h->nreal);
}

/* One radius (R in Mpc, sigma float): filter, c2r, normalise, sigma_grid,
 * barrier, strict max.  Returns seconds; *sigma_grid receives GetMeanSigma. */
double d2z_cpu_radius(void* hv, float R, float sigma, float sigma_min, float erfizeta, float A,
                      double* sigma_grid) {
  d2z_cpu* h = (d2z_cpu*)hv;
  const int N = h->N;
  double t0 = now();
  const float dk = (float)(2. * M_PI / h->box);
  const float kmax = dk * N;
#pragma omp parallel for num_threads(h->nthreads) schedule(static)
  for (int i = 0; i < N; ++i) {
    float kx = i * dk;
    if (i >= N / 2) kx = kx - kmax;
    float kx2 = kx * kx;
    for (int j = 0; j < N; ++j) {
      float ky = j * dk;
      if (j >= N / 2) ky = ky - kmax;
      float ky2 = ky * ky;
      for (int k = 0; k <= N / 2; ++k) {
        float kz = k * dk;
        if (k >= N / 2) kz = kz - kmax;
        float kz2 = kz * kz;
        long index = (long)i * N * (N + 2) + (long)j * (N + 2) + 2 * k;
        float rk = sqrtf(kx2 + ky2 + kz2);
        float rkR = rk * R;
        float num = sinf(rkR) - rkR * cosf(rkR);
        float W = (float)(3. * (double)num / pow((double)rkR, 3));
        if (rk != 0) {
          h->smooth[index] = h->work[index] * W;
          h->smooth[index + 1] = h->work[index + 1] * W;
        } else {
          h->smooth[index] = 0;
          h->smooth[index + 1] = 0;
        }
      }
    }
  }
  axis_fft(h, h->smooth, 0, +1);
  axis_fft(h, h->smooth, 1, +1);
  rows_c2r(h, h->smooth);
  const float norm = powf((float)N, 3.0f);
  double s2 = 0.0;
#pragma omp parallel for num_threads(h->nthreads) schedule(static) reduction(+ : s2)
  for (long row = 0; row < (long)N * N; ++row) {
    float* r = h->smooth + row * (N + 2);
    for (int k = 0; k < N + 2; ++k) r[k] /= norm;
    for (int k = 0; k < N; ++k) s2 += (double)(r[k] * r[k]); /* float product, double sum */
  }
  *sigma_grid = sqrt(s2 / pow((double)(float)N, 3));
  const double Bd = sqrt(2. * (double)(sigma_min * sigma_min - sigma * sigma)) * (double)erfizeta;
  const double Ad = (double)A;
#pragma omp parallel for num_threads(h->nthreads) schedule(static)
  for (long row = 0; row < (long)N * N; ++row) {
    const float* r = h->smooth + row * (N + 2);
    float* zr = h->zreion + row * (N + 2);
    for (int k = 0; k < N; ++k) {
      double zd = Ad * ((double)r[k] + Bd);
      float z = (float)(zd - 1.0);
      if (z > zr[k]) zr[k] = z;
    }
  }
  return now() - t0;
}

/* dense copies out */
void d2z_cpu_get_zreion(void* hv, float* out) {
  d2z_cpu* h = (d2z_cpu*)hv;
  const int N = h->N;
  for (long row = 0; row < (long)N * N; ++row) memcpy(out + row * N, h->zreion + row * (N + 2), sizeof(float) * N);
}

void d2z_cpu_get_smooth(void* hv, float* out) {
  d2z_cpu* h = (d2z_cpu*)hv;
  const int N = h->N;
  for (long row = 0; row < (long)N * N; ++row) memcpy(out + row * N, h->smooth + row * (N + 2), sizeof(float) * N);
}

/* complex half spectrum [N][N][N/2+1] of the last forward transform */
void d2z_cpu_get_spectrum(void* hv, float* out) {
  d2z_cpu* h = (d2z_cpu*)hv;
  memcpy(out, h->work, sizeof(float) * h->nreal);
}

/* standalone 1D transform (for the FFT known-answer tests) */
int d2z_cpu_fft1d(int n, int sign, float* data) {
  plan1d p;
  if (plan_init(&p, n)) return -1;
  cpx* s = (cpx*)malloc(sizeof(cpx) * n);
  fft1d(&p, sign, (cpx*)data, s);
  free(s);
  plan_free(&p);
  return 0;
}
