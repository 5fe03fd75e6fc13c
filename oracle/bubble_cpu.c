/* bubble_cpu.c -- C restatement of pr-calc's history correction of zreion
 * (marcelo-alvarez/pr-calc, c++/src/bubblehistory_d2z.C:44-343 and
 * c++/src/tables.C:9-38), one MPI rank.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the library's skewer, corrected
 * history and remap paths.  The product never links or runs it.
 *
 * Follows:
 *   bubblehistory_d2z.C:104-125  zcur where the float history crosses x (strict
 *                                bracket; no bracket keeps the previous zcur,
 *                                the first one starting from 49.9)
 *   bubblehistory_d2z.C:127-233  skewers: start cells with zreion > zcur and
 *                                j, k < N-2, walks along +y and +z to the first
 *                                zreion < zcur or back to the start, distance
 *                                exp(clow + (chigh - clow) frac) CellSize
 *   bubblehistory_d2z.C:236-240  mfp = sum / (2 nskewers)
 *   bubblehistory_d2z.C:44-80    GetCHistory, float arithmetic
 *   bubblehistory_d2z.C:246-266  BubbleMFP interpolation
 *   tables.C:27-38               Float2Float
 *   bubblehistory_d2z.C:270-343  zreioncorrect, CorrectHistory
 * Deliberate differences, shared with the library: random numbers from
 * Philox4x32-10 keyed by (seed) with counter (skewer, fraction, draw, 0), start
 * coordinates from words 0-2, the two distance fractions from counter (skewer,
 * fraction, 0xffffffff, 0) words 0-1, u = (w >> 8) 2^-24 (the reference uses
 * rand()); distances, their logs and the per-fraction sums in double (the
 * reference: float); at most 16384 start draws per skewer (the reference loops
 * without bound); the last interval of the fcoll table and bins outside the
 * history are clamped (the reference reads past its arrays there).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#define NHISTORY 100000
#define NCHISTORY 10000
#define NX_MFP 1000
#define NMFP 200
#define MAX_DRAWS 16384
#define H_INI 49.9
#define H_FIN 0.01
#define X_INI 0.01
#define X_FIN 0.999

static void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = (uint32_t)p1;
    c[2] = n2;
    c[3] = (uint32_t)p0;
  }
}

static float u24(uint32_t w) { return (float)(w >> 8) * 0x1p-24f; }

static int cell_of(float u, int n) {
  int c = (int)(u * (float)n);
  if (c >= n) c -= n;
  if (c < 0) c += n;
  return c;
}

static float dzh(void) { return (float)((H_FIN - H_INI) / (NHISTORY - 1)); }
static float dzc(void) { return (float)((H_FIN - H_INI) / (NCHISTORY - 1)); }
static float dxm(void) { return (float)((X_FIN - X_INI) / (NX_MFP - 1)); }

void bub_zcur(const float* history, float* zcur, float* xs) {
  const float dz = dzh(), dx = dxm();
  float prev = (float)H_INI;
  for (int ix = 0; ix < NX_MFP; ++ix) {
    const float x = (float)(X_INI + (double)((float)ix * dx));
    float z = prev;
    if (x < history[0]) {
      z = (float)H_INI;
    } else if (x > history[NHISTORY - 1]) {
      z = (float)H_FIN;
    } else {
      for (int i = 0; i < NHISTORY - 1; ++i)
        if (x > history[i] && x < history[i + 1]) {
          const float f = (x - history[i]) / (history[i + 1] - history[i]);
          z = (float)((double)(1.0f - f) * (H_INI + (double)((float)i * dz)) +
                      (double)f * (H_INI + (double)((float)(i + 1) * dz)));
          break;
        }
    }
    zcur[ix] = prev = z;
    xs[ix] = x;
  }
}

static double distance(int m, float frac, double cell) {
  const double clow = m > 1 ? log((double)m - 1.0) : log(0.5);
  const double chigh = log((double)m);
  return exp(clow + (chigh - clow) * (double)frac) * cell;
}

/* zre: dense N^3, index (i N + j) N + k.  Returns 0, or -1 when a skewer found
 * no start cell.  sums[NX_MFP] (double), counts[NX_MFP * NMFP]. */
int bub_skewers(const float* zre, int N, float box, int nskewers, uint64_t seed, const float* zcur, double* sums,
                uint32_t* counts) {
  const float cellf = box / (float)N;
  const double cell = (double)cellf;
  const float lmin = logf(cellf), lmax = (float)log(1.0 * (double)box);
  const float dmfp = (lmax - lmin) / (float)NMFP;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const long long NN = (long long)N * N;
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : failed)
  for (int ix = 0; ix < NX_MFP; ++ix) {
    const float zc = zcur[ix];
    double sum = 0.0;
    for (int s = 0; s < nskewers && !failed; ++s) {
      int i0 = 0, j0 = 0, kk0 = 0, found = 0;
      for (uint32_t a = 0; a < MAX_DRAWS && !found; ++a) {
        uint32_t c[4] = {(uint32_t)s, (uint32_t)ix, a, 0u};
        philox(c, k0, k1);
        i0 = cell_of(u24(c[0]), N);
        j0 = cell_of(u24(c[1]), N);
        kk0 = cell_of(u24(c[2]), N);
        found = zre[i0 * NN + (long long)j0 * N + kk0] > zc && j0 < N - 2 && kk0 < N - 2;
      }
      if (!found) {
        failed = 1;
        break;
      }
      uint32_t c[4] = {(uint32_t)s, (uint32_t)ix, 0xffffffffu, 0u};
      philox(c, k0, k1);
      double d[2];
      for (int dir = 0; dir < 2; ++dir) {
        int j = j0, k = kk0, m = 0;
        for (;;) {
          if (dir == 0) {
            j = (j + 1 == N) ? 0 : j + 1;
          } else {
            k = (k + 1 == N) ? 0 : k + 1;
          }
          ++m;
          const float zv = zre[i0 * NN + (long long)j * N + k];
          if (zv < zc || (dir == 0 ? j == j0 : k == kk0)) break;
        }
        d[dir] = distance(m, u24(c[dir]), cell);
      }
      sum += d[0] + d[1];
      for (int q = 0; q < 2; ++q) {
        const double ld = log(d[q]);
        const int b = (int)((ld - (double)lmin) / (double)dmfp);
        if (b >= 0 && b < NMFP && ld < (double)lmax && ld > (double)lmin) counts[ix * NMFP + b]++;
      }
    }
    sums[ix] = sum;
  }
  return failed ? -1 : 0;
}

static float table_value(const float* t, int n, float tmin, float tmax, float v) {
  if (v <= tmin) return t[0];
  if (v >= tmax) return t[n - 1];
  const float delta = (tmax - tmin) / (float)n;
  const int bin = (int)((v - tmin) / delta);
  if (bin >= n - 1) return t[n - 1];
  const float f = (v - (tmin + (float)bin * delta)) / delta;
  return (1.0f - f) * t[bin] + f * t[bin + 1];
}

static float mfp_at(const float* mfp, float x, float dx) {
  const int itab = (int)(((double)x - X_INI) / (double)dx);
  if (itab < 0) return mfp[0];
  if (itab >= NX_MFP - 1) return mfp[NX_MFP - 1];
  const float f = (float)(((double)(x - (float)itab * dx) - X_INI) / (double)dx);
  return (1.0f - f) * mfp[itab] + f * mfp[itab + 1];
}

float bub_table_value(const float* t, int n, float tmin, float tmax, float v) {
  return table_value(t, n, tmin, tmax, v);
}

float bub_mfp_at(const float* mfp, float x) { return mfp_at(mfp, x, dxm()); }

void bub_chistory(const float* fcoll, int n, float fmin, float fmax, const float* mfp, float rmfp, float zeta,
                  float* ch) {
  const float dz = dzc(), dx = dxm();
  ch[0] = 0.0f;
  for (int i = 1; i < NCHISTORY; ++i) {
    const float x1 = ch[i - 1];
    const float z1 = (float)(H_INI + (double)((float)i * dz));
    const float z2 = (float)(H_INI + (double)((float)(i + 1) * dz));
    const float df = table_value(fcoll, n, fmin, fmax, z2) - table_value(fcoll, n, fmin, fmax, z1);
    const float rb = mfp_at(mfp, x1, dx);
    float v = ch[i - 1] + df / (1.0f + rb / rmfp) * zeta;
    if (v >= 1.0f) v = 1.0f;
    ch[i] = v;
  }
}

/* zvz[NHISTORY] as CorrectHistory builds it, before (filled = 0) or after the
 * backward fill; returns -1 when a history value has no bracketing interval */
int bub_zvz(const float* history, const float* ch, float* zvz, int filled) {
  const float dz = dzc();
  for (int i = 0; i < NHISTORY; ++i) zvz[i] = -1.0f;
  for (int i = 1; i < NHISTORY; ++i) {
    const float x = history[i];
    if (x == 0.0f) continue;
    if (x < ch[0]) {
      zvz[i] = (float)H_INI;
    } else if (x > ch[NCHISTORY - 1]) {
      zvz[i] = (float)H_FIN;
    } else {
      int q = 0;
      for (; q < NCHISTORY - 1; ++q)
        if (x >= ch[q] && x <= ch[q + 1]) break;
      if (q == NCHISTORY - 1) return -1;
      const float f = (x - ch[q]) / (ch[q + 1] - ch[q]);
      zvz[i] = (float)((double)(1.0f - f) * (H_INI + (double)((float)q * dz)) +
                       (double)f * (H_INI + (double)((float)(q + 1) * dz)));
    }
  }
  if (filled)
    for (int i = NHISTORY - 2; i >= 0; --i)
      if (zvz[i] < 0.0f) zvz[i] = zvz[i + 1];
  return 0;
}

void bub_remap(float* zre, long long n, const float* zvz) {
  const double dz = (double)dzh();
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < n; ++c) {
    const double q = ((double)zre[c] - H_INI) / dz;
    const int b = q >= (double)NHISTORY ? NHISTORY - 1 : (q > 0.0 ? (int)q : 0);
    zre[c] = zvz[b];
  }
}
