"""FP64 CPU restatement of pr-calc's ``delta2zreion`` path (the oracle).

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the host
driver, the Python binding) imports, links or runs this file.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may use
it, and only as the checker.

What it restates (reference = marcelo-alvarez/pr-calc, paths relative to the
repository root):

* scalar set-up with the reference's float/double mixture reproduced step by
  step, so the float constants it produces are the ones the reference would
  print:  ``growth`` (c++/src/cosmology.C:174-212), ``transeh`` (:290-309),
  ``sgm2k`` (:274-288), ``cdmehint`` (:214-244), ``normalize`` (:246-271),
  ``inverse_erfc`` (c++/src/math.C:23-61, GSL Brent solver restated),
  the radius schedule (c++/src/delta2zreion_d2z.C:36-38,55), the barrier
  constants of ``d2z`` (:206-216) and the growth factors of
  c++/src/main_d2z.C:57-64;
* the field path in FP64:  forward r2c (main_d2z.C:67), per radius the top-hat
  filter with k=0 zeroed (delta2zreion_d2z.C:58-89), c2r + 1/N^3
  (:92-93), sigma_grid (c++/src/arrayoperations.C:4-35), the barrier and the
  strict running maximum (delta2zreion_d2z.C:106-146);
* ``GetHistory`` and ``tau`` (c++/src/bubblehistory_d2z.C:5-41,
  c++/src/cosmology.C:52-83) with the reference's float accumulation;
* the linear displacement field of ``Delta2Velocity``
  (c++/src/delta2velocity_map.C:8-97) and the binned power spectrum of
  ``PowerSpectrum`` (c++/src/powerspectrum_lmb.C:8-139), downstream uses of
  the same forward transform.

Third-party arithmetic the reference calls, restated here:
* GSL ``gsl_integration_qags`` -> ``scipy.integrate.quad`` (both are QUADPACK
  QAGS: 21-point Gauss-Kronrod with Wynn epsilon extrapolation), same limits,
  epsabs=0, epsrel=1e-7, limit=1000.  GSL version unpinned by the reference
  (its NERSC settings name gsl-2.7).
* GSL ``gsl_root_fsolver_brent`` + ``gsl_root_test_interval`` -> ``_brent_gsl``
  below, Brent's method (R. P. Brent, Algorithms for Minimization without
  Derivatives, 1973, ch. 4) in the bracket-keeping form GSL uses.
* FFTW3 (single precision, FFTW_ESTIMATE) -> ``numpy.fft`` (pocketfft) in FP64.

Pinning (tests/test_oracle_*.py): FFT by brute-force DFT and plane waves,
W(x) by closed-form values and the Taylor series, growth by the exact
Einstein-de Sitter answer and by the integral solution of the linear growth
equation, transeh by its k->0 limit and the EH98 sound horizon, sigma(R) by the
sigma8 identity, inverse_erfc by scipy.special.erfcinv and the GSL bracket
bound, the barrier by its known answer at delta_R = 0, the running maximum by
its zero-floor invariant and GetHistory/tau by brute force on small grids.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy import integrate

f32 = np.float32
f64 = np.float64

NHISTORY = 100000            # c++/src/allvars_d2z.h:1
HISTORY_INITIAL = 49.9       # c++/src/allvars_d2z.h:2
HISTORY_FINAL = 0.01         # c++/src/allvars_d2z.h:3
Z_HIGH = 20.0                # c++/src/main_d2z.C:62
DELTA_C = 1.686              # c++/src/delta2zreion_d2z.C:209


# --------------------------------------------------------------------------
# parameters: parameter-file tags (c++/src/io_d2z.C:33-87) + command line
# (c++/src/commandline_d2z.C:47-52)
# --------------------------------------------------------------------------
@dataclass
class D2ZParams:
    Omegam: float = 0.31
    Omegab: float = 0.049
    Omegal: float = 0.69
    h: float = 0.68
    ns: float = 0.965
    sigma8: float = 0.81
    w: float = -1.0
    BoxSize: float = 200.0       # Mpc/h, as written in the parameter file
    zInit: float = 0.0
    Rmin: float = 1.0            # cells
    N: int = 128
    Nscales: int = 32
    RecordBubbleSizes: int = 0
    Rmax: float = 50.0           # Mpc/h (command line -r)
    zeta: float = 40.0           # command line -z
    M_min: float = 1.4706e8      # Msun (command line -m)


# --------------------------------------------------------------------------
# cosmology.C restated
# --------------------------------------------------------------------------
def growth(z, omegam, omegal, w):
    """D(z), Carroll-Press-Turner fit with the Ma et al. w correction.

    c++/src/cosmology.C:174-212.  Float variables are float32 here; terms the
    C++ promotes to double (``2.5*``, ``pow(float,double)``, the w-fit
    polynomials) are evaluated in float64 and stored back to float32, as the
    reference does.
    """
    z, omegam, omegal, w = f32(z), f32(omegam), f32(omegal), f32(w)
    x = f32(f32(1) + z)
    x2 = f32(x * x)
    x3 = f32(f32(x * x) * x)
    x3w = f32(np.power(x3, w))
    curv = f32(f32(f32(1) - omegam) - omegal)
    omega = f32(f32(omegam * x3) / f32(f32(f32(omegam * x3) + f32(curv * x2)) + omegal))
    lam = f32(f32(f32(omegal * x3) * x3w)
              / f32(f32(f32(omegam * x3) + f32(curv * x2)) + f32(f32(omegal * x3) * x3w)))

    def cpt(om, la):
        # g = 2.5*om/(pow(om,4./7.)-la+(1+om/2)*(1+la/70)); double expression
        tail = f32(f32(f32(1) + f32(om / f32(2))) * f32(f32(1) + f32(la / f32(70))))
        return f32(2.5 * f64(om) / (f64(om) ** (4.0 / 7.0) - f64(la) + f64(tail)))

    g = cpt(omega, lam)
    g0 = cpt(omegam, omegal)
    wd = f64(w)
    a1 = 0.255 + 0.305 * wd + 0.0027 / wd
    a2 = 0.366 + 0.266 * wd - 0.0700 / wd
    # ( 1-omega ) is int - float, i.e. a float difference
    t = f32(-a1 * f64(f32(f32(1) - omega)) - a2 * f64(f32(np.log(omega))))
    t0 = f32(-a1 * f64(f32(f32(1) - omegam)) - a2 * f64(f32(np.log(omegam))))
    g = f32(g * f32(np.power(f32(-w), t)))
    g0 = f32(g0 * f32(np.power(f32(-w), t0)))
    return f32(f32(g / x) / g0)


def deltac(z, omegam, omegal, w):
    """Linear collapse threshold over D(z), cosmology.C:141-172: the open-universe
    (Omega_Lambda = 0, Omega_m < 1) or the flat fitting form of the threshold,
    float variables with the double sub-expressions of the C++ (pow with a double
    or int exponent, the 1.5/3/20/0.0123 factors)."""
    z, omegam, omegal = f32(z), f32(omegam), f32(omegal)
    pi = f32(3.14159265358979)
    x = f32(f32(1) + z)
    x2 = f32(x * x)
    x3 = f32(f32(x * x) * x)
    curv = f32(f32(f32(1) - omegam) - omegal)
    omega = f32(f32(omegam * x3) / f32(f32(f32(omegam * x3) + f32(curv * x2)) + omegal))
    u = f32(f32(f32(2) / omega) - f32(1))
    eta = f32(math.log(f64(u) + math.sqrt(f64(u) ** 2 - 1)))
    if omegam < 1 and omegal == 0:
        sh, ch = f32(np.sinh(eta)), f32(np.cosh(eta))
        a = f32(f32(f32(3) * sh) * f32(sh - eta))
        dc = f32(1.5 * (f64(a) / f64(f32(ch - f32(1))) ** 2 - 2)
                 * (1 + f64(f32(f32(f32(2) * pi) / f32(sh - eta))) ** (2.0 / 3.0)))
    else:
        dc = f32(3 * f64(f32(f32(12) * pi)) ** (2.0 / 3.0) / 20 * (1 + 0.0123 * f64(f32(np.log10(omega)))))
    return f32(dc / growth(z, omegam, omegal, w))


def fcoll(z, omegam, omegal, w, sigma_min):
    """Collapsed fraction erfc(deltac(z) / sqrt(2) / sigma_min), cosmology.C:311-324
    (the fourth column of WriteHistory, io_d2z.C:284-314)."""
    arg = f32(f64(deltac(z, omegam, omegal, w)) / math.sqrt(2.0) / f64(f32(sigma_min)))
    return f32(math.erfc(f64(arg)))


class _EH:
    """Common block of cosmology.C:14 (omega0, baryon0, hubble0, an0, pnorm0)."""

    def __init__(self, omega, baryon, hubble, an, pnorm):
        self.omega0 = f64(f32(omega))
        self.baryon0 = f64(f32(baryon))
        self.hubble0 = f64(f32(hubble))
        self.an0 = f64(f32(an))
        self.pnorm0 = f64(pnorm)

    def transeh(self, k):
        """Eisenstein & Hu 1998 no-wiggle transfer function, cosmology.C:290-309.
        k in 1/Mpc."""
        om, ob, hh = self.omega0, self.baryon0, self.hubble0
        alpha = (1 - 0.328 * np.log(431 * om * hh * hh) * (ob / om)
                 + 0.38 * np.log(22.3 * om * hh * hh) * (ob / om) ** 2)
        sound = 44.5 * np.log(9.83 / (om * hh * hh)) / np.sqrt(1 + 10 * (ob * hh * hh) ** 0.75)
        shape = om * hh * (alpha + (1 - alpha) / (1 + (0.43 * k * sound) ** 4))
        q = (k / hh) * (2.728 / 2.7) ** 2 / shape
        L = np.log(2 * math.e + 1.8 * q)
        return L / (L + (14.2 + 731 / (1 + 62.5 * q)) * q * q)

    def sgm2k(self, lk, R):
        """Integrand of sigma^2 in ln k, cosmology.C:274-288."""
        k = np.exp(lk)
        kR = k * R
        Pin = k ** (3 + self.an0)
        T = self.transeh(k)
        W = 3 * (np.sin(kR) - kR * np.cos(kR)) / kR ** 3
        return Pin * T * T * W * W * self.pnorm0 ** 2

    def integral(self, R):
        """gsl_integration_qags(sgm2k, log(1e-5/R), log(1e5/R), 0, 1e-7, 1000)."""
        val, _err = integrate.quad(self.sgm2k, math.log(1e-5 / R), math.log(1e5 / R),
                                   args=(R,), epsabs=0.0, epsrel=1e-7, limit=1000)
        return val


def normalize(omega, baryon, hubble, an, sigma8):
    """pnorm so that sigma(8 Mpc/h) = sigma8, cosmology.C:246-271 (returns float)."""
    eh = _EH(omega, baryon, hubble, an, 1.0)
    R = f64(f32(f32(8) / f32(hubble)))          # R = 8/hubble (float division)
    s2 = eh.integral(R)
    return f32(f64(f32(sigma8)) / math.sqrt(s2))


def cdmehint(omega, baryon, hubble, an, pnorm, mass):
    """sigma^2(M) in float, cosmology.C:214-244."""
    eh = _EH(omega, baryon, hubble, an, f64(f32(pnorm)))
    rhoc = f32(2.775e11)
    pi = f32(3.141592654)
    fourpi = f32(f32(4) * pi)
    dens = f64(f32(f32(f32(f32(omega) * rhoc) * f32(hubble)) * f32(hubble)))
    num = f32(f32(f32(3) * f32(mass)) / fourpi)
    R = (f64(num) / dens) ** (1.0 / 3.0)
    return f32(eh.integral(R))


def radius_mass(R, omegam, h):
    """mass = 4*M_PI/3*rho0*R^3 with float rho0, delta2zreion_d2z.C:15,97."""
    rho0 = f32(2.775e11 * f64(f32(omegam)) * f64(f32(h)) * f64(f32(h)))
    R = f64(f32(R))
    return f32(4 * math.pi / 3 * f64(rho0) * R * R * R)


# --------------------------------------------------------------------------
# math.C restated: erfc^-1 by Brent's method with GSL's bracket semantics
# --------------------------------------------------------------------------
def _brent_gsl(f, x_lower, x_upper, epsabs, epsrel, max_iter):
    """Brent's root bracketing method, iterated until the bracket test
    |x_hi - x_lo| < epsabs + epsrel*min(|x_lo|,|x_hi|) (bracket of one sign) passes.
    Returns (x_lo, x_hi, converged).  The state update follows Brent (1973):
    keep (b, c) bracketing the root with |f(b)| <= |f(c)|, try inverse quadratic
    (or secant) interpolation, fall back to bisection when the step is not
    inside 3/4 of the bracket or does not shrink fast enough."""
    eps = np.finfo(np.float64).eps
    a, b = x_lower, x_upper
    fa, fb = f(a), f(b)
    c, fc = b, fb
    d = e = x_upper - x_lower
    lo, hi = x_lower, x_upper
    for _ in range(max_iter):
        ac_equal = False
        if (fb < 0 and fc < 0) or (fb > 0 and fc > 0):
            ac_equal = True
            c, fc = a, fa
            d = e = b - a
        if abs(fc) < abs(fb):
            ac_equal = True
            a, b, c = b, c, b
            fa, fb, fc = fb, fc, fb
        tol = 0.5 * eps * abs(b)
        m = 0.5 * (c - b)
        if fb == 0:
            return b, b, True
        if abs(m) <= tol:
            lo, hi = (b, c) if b < c else (c, b)
            return lo, hi, True
        if abs(e) < tol or abs(fa) <= abs(fb):
            d = e = m
        else:
            s = fb / fa
            if ac_equal:
                p = 2 * m * s
                q = 1 - s
            else:
                q = fa / fc
                r = fb / fc
                p = s * (2 * m * q * (q - r) - (b - a) * (r - 1))
                q = (q - 1) * (r - 1) * (s - 1)
            if p > 0:
                q = -q
            else:
                p = -p
            if 2 * p < min(3 * m * q - abs(tol * q), abs(e * q)):
                e = d
                d = p / q
            else:
                d = e = m
        a, fa = b, fb
        b += d if abs(d) > tol else (tol if m > 0 else -tol)
        fb = f(b)
        if (fb < 0 and fc < 0) or (fb > 0 and fc > 0):
            c = a
        lo, hi = (b, c) if b < c else (c, b)
        # gsl_root_test_interval(lo, hi, epsabs, epsrel)
        if (lo > 0 and hi > 0) or (lo < 0 and hi < 0):
            min_abs = min(abs(lo), abs(hi))
        else:
            min_abs = 0.0
        if abs(hi - lo) < epsabs + epsrel * min_abs:
            return lo, hi, True
    return lo, hi, False


def inverse_erfc(x):
    """math.C:23-61: Brent on erfc(t) - x over [0, 100], test_interval(0, 1e-4),
    100 iterations, returns the LOWER end of the final bracket as float.
    The argument arrives as float (``inverse_erfc(float x)``)."""
    a = f64(f32(x))
    lo, _hi, _ok = _brent_gsl(lambda t: math.erfc(t) - a, 0.0, 100.0, 0.0, 1e-4, 100)
    return f32(lo)


# --------------------------------------------------------------------------
# scalar set-up of delta2zreion (delta2zreion_d2z.C:12-38, main_d2z.C:29,57-64)
# --------------------------------------------------------------------------
@dataclass
class Scalars:
    DInit: np.float32
    DHigh: np.float32
    A: np.float32                   # (1+zHigh)*DHigh/deltac, float
    pnorm: np.float32
    erfizeta: np.float32
    sigma_min: np.float32
    box_mpc: np.float32             # BoxSize/h (float)
    R: np.ndarray = field(default=None)       # float32[Nscales], Mpc
    sigma: np.ndarray = field(default=None)   # float32[Nscales]
    B: np.ndarray = field(default=None)       # float64[Nscales] = sqrt(2(smin^2-s^2))*erfizeta


def radius_schedule(p: D2ZParams):
    """R_r = 10^(log10 R0 + r dScale), delta2zreion_d2z.C:36-38,55 (float)."""
    box = f32(f32(p.BoxSize) / f32(p.h))
    Rinitial = f32(f32(box * f32(p.Rmin)) / f32(p.N))
    Rfinal = f32(f32(p.Rmax) / f32(p.h))
    # log10 of a float, correctly rounded (what glibc's log10f returns)
    l0 = f32(math.log10(f64(Rinitial)))
    l1 = f32(math.log10(f64(Rfinal)))
    dScale = f32(f32(l1 - l0) / f32(p.Nscales - 1))
    out = np.empty(p.Nscales, dtype=np.float32)
    for r in range(p.Nscales):
        out[r] = f32(10.0 ** f64(f32(l0 + f32(f32(r) * dScale))))
    return out


def barrier_B(sigma_min, sigma, erfizeta):
    """sqrt(2.*(sigma_min*sigma_min-sigma*sigma))*erfizeta in double, with the
    float difference of squares, delta2zreion_d2z.C:211-212."""
    d = f32(f32(sigma_min * sigma_min) - f32(f32(sigma) * f32(sigma)))
    with np.errstate(invalid="ignore"):
        return np.sqrt(2.0 * f64(d)) * f64(erfizeta)


def setup_scalars(p: D2ZParams) -> Scalars:
    DInit = growth(p.zInit, p.Omegam, p.Omegal, p.w)
    DHigh = growth(Z_HIGH, p.Omegam, p.Omegal, p.w)
    A = f32(f32(f32(f32(1) + f32(Z_HIGH)) * DHigh) / f32(DELTA_C))
    pnorm = normalize(p.Omegam, p.Omegab, p.h, p.ns, p.sigma8)
    erfizeta = inverse_erfc(1.0 / f64(f32(p.zeta)))
    smin = f32(np.sqrt(cdmehint(p.Omegam, p.Omegab, p.h, p.ns, pnorm, f32(p.M_min))))
    R = radius_schedule(p)
    sig = np.empty_like(R)
    for r, Rr in enumerate(R):
        m = radius_mass(Rr, p.Omegam, p.h)
        sig[r] = f32(np.sqrt(cdmehint(p.Omegam, p.Omegab, p.h, p.ns, pnorm, m)))
    B = np.array([barrier_B(smin, s, erfizeta) for s in sig], dtype=np.float64)
    return Scalars(DInit=DInit, DHigh=DHigh, A=A, pnorm=pnorm, erfizeta=erfizeta,
                   sigma_min=smin, box_mpc=f32(f32(p.BoxSize) / f32(p.h)), R=R, sigma=sig, B=B)


# --------------------------------------------------------------------------
# field path (FP64)
# --------------------------------------------------------------------------
def tophat_W(x):
    """W(x) = 3 (sin x - x cos x)/x^3, evaluated with its Taylor series
    1 - x^2/10 + x^4/280 - x^6/15120 + ... (term ratio -x^2/(2n(2n+3)), through
    x^12) for x < 0.5 (the reference evaluates the closed form in float,
    delta2zreion_d2z.C:76-77)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.empty_like(x)
    small = x < 0.5
    xs = x[small]
    x2 = xs * xs
    out[small] = 1 - x2 / 10 * (1 - x2 / 28 * (1 - x2 / 54 * (1 - x2 / 88 * (1 - x2 / 130 * (1 - x2 / 180)))))
    xl = x[~small]
    out[~small] = 3 * (np.sin(xl) - xl * np.cos(xl)) / xl ** 3
    return out


def kmag_grid(N, box_mpc):
    """|k| on the half-spectrum grid [kx][ky][kz<=N/2]; indices >= N/2 map to
    negative frequencies (delta2zreion_d2z.C:59-71)."""
    dk = 2 * math.pi / f64(box_mpc)
    i = np.arange(N)
    kk = np.where(i >= N // 2, i - N, i).astype(np.float64) * dk
    kz = np.arange(N // 2 + 1).astype(np.float64) * dk
    kz = np.where(np.arange(N // 2 + 1) >= N // 2, kz - N * dk, kz)
    return np.sqrt(kk[:, None, None] ** 2 + kk[None, :, None] ** 2 + kz[None, None, :] ** 2)


def forward(delta, DInit, workers=None):
    """delta /= DInit then unnormalised r2c (main_d2z.C:57-59,67).  With
    ``workers`` the transform runs on scipy.fft's threaded FFT (large boxes)."""
    x = np.asarray(delta, dtype=np.float64) / f64(DInit)
    if workers:
        import scipy.fft
        return scipy.fft.rfftn(x, workers=workers)
    return np.fft.rfftn(x)


def smooth(dhat, kmag, R, workers=None):
    """Filtered real field delta_R = c2r(W(kR) dhat)/N^3 with k=0 zeroed
    (delta2zreion_d2z.C:58-93)."""
    N = dhat.shape[0]
    W = tophat_W(kmag * f64(R))
    W.flat[0] = 0.0
    if workers:
        import scipy.fft
        return scipy.fft.irfftn(dhat * W, s=(N, N, N), axes=(0, 1, 2), workers=workers)
    return np.fft.irfftn(dhat * W, s=(N, N, N), axes=(0, 1, 2))


def displacement(delta, box, axis, workers=None):
    """Linear displacement component psi_a = c2r(i k_a / k^2 * r2c(delta)) / N^3
    with k = n 2 pi / box (signed n: n >= N/2 -> n - N), k = 0 and the
    component's own Nyquist plane zeroed: Delta2Velocity's xdv_a
    (c++/src/delta2velocity_map.C:8-97; its two NormalizeArray(.., 1.5) calls,
    :24,90-92, give 1/N^3 in all)."""
    N = delta.shape[0]
    dk = 2.0 * math.pi / f64(box)
    n = np.fft.fftfreq(N, 1.0 / N)                  # 0..N/2-1, -N/2..-1
    nz = np.arange(N // 2 + 1, dtype=np.float64)     # 0..N/2
    kx = (dk * n)[:, None, None]
    ky = (dk * n)[None, :, None]
    kz = (dk * nz)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    ka = np.broadcast_to((kx, ky, kz)[axis], k2.shape)
    with np.errstate(divide="ignore", invalid="ignore"):
        m = np.where(k2 > 0, 1j * ka / np.where(k2 > 0, k2, 1.0), 0.0)
    nyq = (np.abs(n) == N // 2)
    if axis == 0:
        m[nyq, :, :] = 0.0
    elif axis == 1:
        m[:, nyq, :] = 0.0
    else:
        m[:, :, N // 2] = 0.0
    if workers:
        import scipy.fft
        dhat = scipy.fft.rfftn(f64(delta), axes=(0, 1, 2), workers=workers)
        return scipy.fft.irfftn(dhat * m, s=(N, N, N), axes=(0, 1, 2), workers=workers)
    return np.fft.irfftn(np.fft.rfftn(f64(delta)) * m, s=(N, N, N), axes=(0, 1, 2))


def power_spectrum(delta, box, nk, kmin, kmax):
    """PowerSpectrum (c++/src/powerspectrum_lmb.C:8-139, transform = 1): modes
    of the half spectrum (kz = 0..N/2) of FFT(delta) N^-1.5 (:37-38), binned by
    bin = (int)((log k - log kmin) / dlnk) (:73) with the reference's float
    set-up (:18-26) and float k (:55-71); per bin the count, (2 pi)^3
    <|delta_k|^2> / d3k, <1/k^2> and <k> (:78-81,108-110; NaN when empty).
    The power itself in FP64."""
    f32 = np.float32
    N = delta.shape[0]
    fk = np.fft.rfftn(f64(delta)) * f64(N) ** -1.5
    dk = f32(2.0 * math.pi / f64(f32(box)))
    kmaxg = f32(dk * f32(N))
    d3k = f32(f64(kmaxg) ** 3)
    lnkmin = np.log(f32(kmin))
    dlnk = f32((np.log(f32(kmax)) - lnkmin) / f32(nk))

    def axis_k(n):
        i = np.arange(n)
        k = (i.astype(f32) * dk).astype(f32)
        return np.where(i >= N // 2, (k - kmaxg).astype(f32), k)

    kxy, kz = axis_k(N), axis_k(N // 2 + 1)
    rk2 = ((kxy * kxy)[:, None, None] + (kxy * kxy)[None, :, None]).astype(f32) + (kz * kz)[None, None, :]
    rk2 = rk2.astype(f32)
    rk = np.sqrt(rk2)
    with np.errstate(divide="ignore", invalid="ignore"):
        u = ((np.log(rk) - lnkmin) / dlnk).astype(f32)
    sel = (u > -1.0) & (u < nk)
    b = np.trunc(u[sel]).astype(np.int64)
    power = (fk.real ** 2 + fk.imag ** 2)[sel]
    n = np.bincount(b, minlength=nk).astype(np.int64)
    with np.errstate(divide="ignore", invalid="ignore"):
        ps = np.bincount(b, power / f64(d3k), minlength=nk) * (2.0 * math.pi) ** 3 / n
        ook2 = np.bincount(b, (f32(1.0) / rk2[sel]).astype(np.float64), minlength=nk) / n
        kmean = np.bincount(b, rk[sel].astype(np.float64), minlength=nk) / n
    return {"ps": ps, "ook2": ook2, "kmean": kmean, "nmodes": n, "u": u}


def sigma_grid(dR):
    """GetMeanSigma: sqrt(sum dR^2 / N^3), arrayoperations.C:4-35."""
    return math.sqrt(float(np.sum(dR * dR)) / dR.size)


def barrier(dR, A, B):
    """z = A*(delta_R + B) - 1 in double, stored float (delta2zreion_d2z.C:206-216)."""
    return (f64(A) * (dR + B) - 1.0)


def delta2zreion(delta, p: D2ZParams, sc: Scalars | None = None, radii=None,
                 want_rbubble=False, return_fp64=False):
    """Full oracle: returns (zreion float32 [N,N,N], sigma_grid[list], rbubble|None).

    ``radii`` optionally restricts the loop to a subset of radius indices (the
    running maximum is then over that subset only)."""
    if sc is None:
        sc = setup_scalars(p)
    N = p.N
    dhat = forward(delta, sc.DInit)
    kmag = kmag_grid(N, sc.box_mpc)
    zre = np.zeros((N, N, N), dtype=np.float64)
    rb = np.zeros((N, N, N), dtype=np.float32) if want_rbubble else None
    sgrid = []
    idx = range(p.Nscales) if radii is None else radii
    for r in idx:
        dR = smooth(dhat, kmag, sc.R[r])
        sgrid.append(sigma_grid(dR))
        z = barrier(dR, sc.A, sc.B[r])
        if not return_fp64:
            z = z.astype(np.float32).astype(np.float64)
        upd = z > zre                      # strict; NaN never updates
        zre = np.where(upd, z, zre)
        if rb is not None:
            rb[upd] = sc.R[r]
    out = zre if return_fp64 else zre.astype(np.float32)
    return out, sgrid, rb


# --------------------------------------------------------------------------
# GetHistory + tau (bubblehistory_d2z.C:5-41, cosmology.C:52-83)
# --------------------------------------------------------------------------
def history_dz():
    """dzhistory=(HISTORY_FINAL-HISTORY_INITIAL)/(NHISTORY-1) stored as float."""
    return f32((HISTORY_FINAL - HISTORY_INITIAL) / (NHISTORY - 1))


def history_bins(zreion):
    """bin=(int)((zcur-HISTORY_INITIAL)/dzhistory) in double, truncated."""
    dz = f64(history_dz())
    q = (np.asarray(zreion, dtype=np.float32).astype(np.float64) - HISTORY_INITIAL) / dz
    b = np.trunc(q)
    ok = (b >= 0) & (b < NHISTORY)
    return b[ok].astype(np.int64)


def history_counts(zreion):
    """Integer counts per history bin (the reference keeps float counts)."""
    return np.bincount(history_bins(zreion), minlength=NHISTORY).astype(np.int64)


def xhii_from_counts(counts, ncells):
    """Cumulative ionized fraction x_HII(z_i), in double from integer counts."""
    return np.cumsum(counts, dtype=np.int64).astype(np.float64) / float(ncells)


def history_float(counts, N):
    """The reference's float history: counts/N^3 then float cumulative sum."""
    h = (counts.astype(np.float64) / float(N) ** 3).astype(np.float32)
    out = np.empty(NHISTORY, dtype=np.float32)
    acc = f32(0)
    for i in range(NHISTORY):
        acc = f32(acc + h[i]) if i else h[0]
        out[i] = acc
    return out


def history_z():
    dz = history_dz()
    z = np.empty(NHISTORY, dtype=np.float32)
    z[0] = f32(HISTORY_INITIAL)
    for i in range(1, NHISTORY):
        z[i] = f32(HISTORY_INITIAL + f64(f32(f32(i) * dz)))
    return z


def tau(za, xa, omegam, omegab, h):
    """Thomson optical depth, cosmology.C:52-83 (float accumulator)."""
    omegam, omegab, h = f32(omegam), f32(omegab), f32(h)
    thompson = f32(6.65e-25)
    YHe = f32(0.25)
    hubble0 = f32(100.0 / 3.086e19 * f64(h))
    ne0 = f32(f64(f32(f32(omegab * h) * h)) * 1.88e-29 / 1.67e-24)
    c = f32(3e10)
    dtau0 = f32(f32(f32(thompson * ne0) * c) / hubble0)
    t = f32(0)
    nz = len(za)
    for i in range(1, nz - 1):
        z = za[i]
        x = xa[i]
        dz = f32(abs(f32(za[i + 1] - za[i])))
        NHe = f32(2) if z < 3 else f32(1)
        zp = f64(f32(f32(1) + z))
        core = f64(f32(x * dtau0)) * zp ** 2 / math.sqrt(f64(omegam) * zp ** 3 + 1 - f64(omegam))
        fac = f32(f32(f32(1) - YHe) + f32(f32(NHe * YHe) / f32(4)))
        t = f32(f64(t) + core * f64(fac) * f64(dz))
    return t
