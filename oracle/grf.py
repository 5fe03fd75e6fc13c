"""Numpy restatement of the synthetic linear density field used as d2z input.

TEST INFRASTRUCTURE ONLY (the checker for the library's GPU field generator
and the source of test inputs).  The product never imports this file.

The reference's own generator (c++/src/grf_ics.C:8-112) draws real-space
normals with srand48((rank+123)*Seed) and Marsaglia's polar method, so its
output depends on the MPI decomposition; it is not reproduced.  This
generator instead keys every cell's normal on (seed, cell index):

* Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
  easy as 1, 2, 3", SC'11), counter = (cell & 0xffffffff, cell >> 32, 0, 0),
  key = (seed & 0xffffffff, seed >> 32);
* Box-Muller on the first two words: u1 = ((w0 >> 8) + 0.5) 2^-24,
  u2 = (w1 >> 8) 2^-24, g = sqrt(-2 ln u1) cos(2 pi u2), evaluated in double and
  rounded to float;
* colouring in k-space as the reference's ConvolveNoise does
  (c++/src/grf_ics.C:34-93): delta_k = noise_k sqrt(P(k)/V_cell), k=0 set to 0,
  unnormalised forward transform and 1/N^3 inverse;
* P(k) = A k^ns T_BBKS(k)^2 (Bardeen, Bond, Kaiser & Szalay 1986, eq. G3) with
  q = k/Gamma, Gamma = Omega_m h exp(-Omega_b (1 + sqrt(2h)/Omega_m))
  (Sugiyama 1995), k in h/Mpc, A fixed by sigma(8 Mpc/h) = sigma8.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import integrate

PHILOX_M0 = np.uint64(0xD2511F53)
PHILOX_M1 = np.uint64(0xCD9E8D57)
PHILOX_W0 = 0x9E3779B9
PHILOX_W1 = 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Philox4x32 with 10 rounds on uint32 arrays; returns four uint32 arrays."""
    c0 = np.asarray(c0, dtype=np.uint64)
    c1 = np.asarray(c1, dtype=np.uint64)
    c2 = np.asarray(c2, dtype=np.uint64)
    c3 = np.asarray(c3, dtype=np.uint64)
    k0 = int(k0) & 0xFFFFFFFF
    k1 = int(k1) & 0xFFFFFFFF
    for rnd in range(10):
        if rnd:
            k0 = (k0 + PHILOX_W0) & 0xFFFFFFFF
            k1 = (k1 + PHILOX_W1) & 0xFFFFFFFF
        p0 = c0 * PHILOX_M0
        p1 = c2 * PHILOX_M1
        n0 = (p1 >> np.uint64(32)) ^ c1 ^ np.uint64(k0)
        n1 = p1 & _MASK
        n2 = (p0 >> np.uint64(32)) ^ c3 ^ np.uint64(k1)
        n3 = p0 & _MASK
        c0, c1, c2, c3 = n0, n1, n2, n3
    return (c0.astype(np.uint32), c1.astype(np.uint32),
            c2.astype(np.uint32), c3.astype(np.uint32))


def cell_normals(seed, start, count):
    """Unit normals for cells [start, start+count) as float32."""
    n = np.arange(start, start + count, dtype=np.uint64)
    w0, w1, _, _ = philox4x32_10(n & _MASK, n >> np.uint64(32), 0 * n, 0 * n,
                                 seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    u1 = ((w0 >> np.uint32(8)).astype(np.float64) + 0.5) * 2.0 ** -24
    u2 = (w1 >> np.uint32(8)).astype(np.float64) * 2.0 ** -24
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)).astype(np.float32)


def bbks_gamma(omegam, omegab, h):
    return omegam * h * math.exp(-omegab * (1.0 + math.sqrt(2.0 * h) / omegam))


def t_bbks(k, gamma):
    """BBKS transfer function, k in h/Mpc."""
    q = np.asarray(k, dtype=np.float64) / gamma
    with np.errstate(divide="ignore", invalid="ignore"):
        t = (np.log(1 + 2.34 * q) / (2.34 * q)
             * (1 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25)
    return np.where(q == 0, 1.0, t)


def _w_tophat(x):
    x = np.asarray(x, dtype=np.float64)
    small = x < 1e-3
    xs = np.where(small, 1.0, x)
    return np.where(small, 1 - x * x / 10, 3 * (np.sin(xs) - xs * np.cos(xs)) / xs ** 3)


def amplitude(omegam=0.31, omegab=0.049, h=0.68, ns=0.965, sigma8=0.81):
    """A such that sigma(8 Mpc/h) = sigma8 for P = A k^ns T_BBKS^2."""
    gamma = bbks_gamma(omegam, omegab, h)

    def f(lk):
        k = math.exp(lk)
        return k ** (3 + ns) * t_bbks(k, gamma) ** 2 * _w_tophat(8.0 * k) ** 2 / (2 * math.pi ** 2)

    s2, _ = integrate.quad(f, math.log(1e-6), math.log(1e4), epsabs=0, epsrel=1e-10, limit=2000)
    return sigma8 ** 2 / s2


def power(k, omegam=0.31, omegab=0.049, h=0.68, ns=0.965, sigma8=0.81):
    A = amplitude(omegam, omegab, h, ns, sigma8)
    gamma = bbks_gamma(omegam, omegab, h)
    k = np.asarray(k, dtype=np.float64)
    return A * k ** ns * t_bbks(k, gamma) ** 2


def white_noise(N, seed):
    return cell_normals(seed, 0, N ** 3).reshape(N, N, N)


def grf(N, box_mpch, seed, omegam=0.31, omegab=0.049, h=0.68, ns=0.965, sigma8=0.81):
    """Dense float32 [N,N,N] density contrast (x slowest), box in Mpc/h."""
    noise = white_noise(N, seed).astype(np.float64)
    nk = np.fft.rfftn(noise)
    dk = 2 * math.pi / box_mpch
    i = np.arange(N)
    kk = np.where(i >= N // 2, i - N, i) * dk
    kz = np.arange(N // 2 + 1) * dk
    k = np.sqrt(kk[:, None, None] ** 2 + kk[None, :, None] ** 2 + kz[None, None, :] ** 2)
    vcell = (box_mpch / N) ** 3
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.sqrt(power(k, omegam, omegab, h, ns, sigma8) / vcell)
    amp.flat[0] = 0.0
    return np.fft.irfftn(nk * amp, s=(N, N, N), axes=(0, 1, 2)).astype(np.float32)
