"""Pins of the oracle (oracle/d2z_fp64.py, oracle/grf.py, oracle/d2z_cpu.c).

The reference ships no golden vectors for this path, so each oracle function
is checked against something independent of its own formula: a closed form, a
limit, a physical identity, a different algorithm or library, or brute force.
"""
import math

import numpy as np
import pytest
from scipy import integrate, optimize, special

import d2z_cpu
import d2z_fp64 as o
import grf

P = o.D2ZParams()


# ------------------------------------------------------------------ FFT use
def test_forward_and_smooth_of_a_plane_wave():
    """A single plane wave is an eigenfunction of the filter: delta_R = W(|k|R) delta,
    and r2c puts N^3/2 at the wave's two modes (numpy pocketfft, unnormalised)."""
    N, box = 16, 20.0
    x = np.arange(N)
    kx, ky = 3, N - 2   # one positive, one negative frequency index
    d = np.cos(2 * np.pi * (kx * x[:, None, None] + ky * x[None, :, None]) / N) * np.ones((1, 1, N))
    dhat = o.forward(d, 1.0)
    assert abs(dhat[kx, ky, 0] - N ** 3 / 2) < 1e-9
    mask = np.ones(dhat.shape, bool)
    mask[kx, ky, 0] = mask[N - kx, N - ky, 0] = False
    assert np.max(np.abs(dhat[mask])) < 1e-9
    kmag = o.kmag_grid(N, box)
    R = 2.3
    k = 2 * np.pi / box * math.hypot(kx, N - ky)
    dR = o.smooth(dhat, kmag, R)
    np.testing.assert_allclose(dR, o.tophat_W(k * R) * d, atol=1e-12)


def test_kgrid_maps_high_indices_to_negative_frequencies():
    N, box = 16, 10.0
    k = o.kmag_grid(N, box)
    dk = 2 * np.pi / box
    assert k[N - 1, 0, 0] == pytest.approx(dk)
    assert k[N // 2, 0, 0] == pytest.approx(N / 2 * dk)   # Nyquist -> -k_N, |k| = N/2 dk
    assert k[0, 0, N // 2] == pytest.approx(N / 2 * dk)


def test_mean_mode_is_zeroed():
    N = 16
    d = np.full((N, N, N), 0.7)
    dR = o.smooth(o.forward(d, 1.0), o.kmag_grid(N, 10.0), 1.0)
    assert np.max(np.abs(dR)) < 1e-12


@pytest.mark.parametrize("n", [8, 12, 16, 48, 64, 96])
def test_cpu_restatement_fft_matches_brute_force_dft(n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    j = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(j, j) / n)
    ref_f = F @ x.astype(np.complex128)
    ref_b = F.conj() @ x.astype(np.complex128)
    assert np.max(np.abs(d2z_cpu.fft1d(x, -1) - ref_f)) < 1e-5 * np.max(np.abs(ref_f))
    assert np.max(np.abs(d2z_cpu.fft1d(x, +1) - ref_b)) < 1e-5 * np.max(np.abs(ref_b))


def test_cpu_restatement_r2c_matches_numpy():
    N = 16
    rng = np.random.default_rng(5)
    d = rng.standard_normal((N, N, N)).astype(np.float32)
    h = d2z_cpu.CpuD2Z(N, 10.0, 2)
    h.forward(d, 1.0)
    spec = h.spectrum()
    h.close()
    ref = np.fft.rfftn(d.astype(np.float64))
    assert np.max(np.abs(spec - ref)) < 2e-5 * np.max(np.abs(ref))


# -------------------------------------------------------------- top-hat W
def test_tophat_closed_form_values():
    assert o.tophat_W(np.array([0.0]))[0] == 1.0
    assert o.tophat_W(np.array([np.pi]))[0] == pytest.approx(3 / np.pi ** 2, rel=1e-14)
    # first zero: tan x = x
    x0 = optimize.brentq(lambda x: math.tan(x) - x, 4.4, 4.6)
    assert abs(o.tophat_W(np.array([x0]))[0]) < 1e-14
    # large x: W ~ -3 cos x / x^2
    x = np.array([1e4 + 0.3])
    assert o.tophat_W(x)[0] == pytest.approx(-3 * np.cos(x[0]) / x[0] ** 2, rel=1e-3)


def test_tophat_series_meets_closed_form():
    x = np.linspace(0.45, 0.55, 11)
    closed = 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
    np.testing.assert_allclose(o.tophat_W(x), closed, rtol=2e-12)
    xs = np.array([1e-4])
    assert o.tophat_W(xs)[0] == pytest.approx(1 - xs[0] ** 2 / 10, rel=1e-15)


# ----------------------------------------------------------------- growth
def _growth_exact(z, om, ol, w):
    """Linear growth of flat wCDM from the ODE in ln a, normalised to D(z=0)=1."""
    def E2(a):
        return om * a ** -3 + ol * a ** (-3 * (1 + w))

    def rhs(lna, y):
        a = math.exp(lna)
        D, dD = y
        dlnE = 0.5 * (-3 * om * a ** -3 - 3 * (1 + w) * ol * a ** (-3 * (1 + w))) / E2(a)
        omz = om * a ** -3 / E2(a)
        return [dD, -(2 + dlnE) * dD + 1.5 * omz * D]

    a0 = 1e-3
    sol = integrate.solve_ivp(rhs, [math.log(a0), 0.0], [a0, a0], rtol=1e-10, atol=1e-14,
                              dense_output=True)
    Dz = sol.sol(math.log(1 / (1 + z)))[0]
    return Dz / sol.sol(0.0)[0]


def test_growth_einstein_de_sitter_exact():
    for z in (0.0, 1.0, 7.0, 20.0):
        assert o.growth(z, 1.0, 0.0, -1.0) == pytest.approx(1 / (1 + z), rel=2e-7)


def test_growth_fit_tracks_the_growth_equation():
    for z in (0.0, 0.5, 2.0, 6.0, 20.0):
        fit = float(o.growth(z, 0.31, 0.69, -1.0))
        assert fit == pytest.approx(_growth_exact(z, 0.31, 0.69, -1.0), rel=0.005)


def test_growth_with_w_matches_where_matter_dominates():
    # cosmology.C:188 keeps "+ omegal" (no (1+z)^(3w) factor) in the Omega_m(z)
    # denominator, so for w != -1 the reference's D(z) departs from the growth
    # equation by several per cent at z < 2; the oracle reproduces that.  Where
    # matter dominates both agree.
    for z in (6.0, 10.0, 20.0):
        fit = float(o.growth(z, 0.31, 0.69, -0.8))
        assert fit == pytest.approx(_growth_exact(z, 0.31, 0.69, -0.8), rel=0.01)
    assert abs(float(o.growth(0.5, 0.31, 0.69, -0.8)) / _growth_exact(0.5, 0.31, 0.69, -0.8) - 1) > 0.03


# ---------------------------------------------------------------- transeh
def test_transeh_limits_and_baryon_suppression():
    eh = o._EH(0.31, 0.049, 0.68, 0.965, 1.0)
    assert eh.transeh(1e-7) == pytest.approx(1.0, abs=1e-5)
    # EH98 eq. 26 sound horizon for omega_m h^2 = 0.1432, omega_b h^2 = 0.02237: ~147 Mpc
    om, ob = 0.1432, 0.02237
    s = 44.5 * math.log(9.83 / om) / math.sqrt(1 + 10 * ob ** 0.75)
    assert s == pytest.approx(147.1, rel=0.03)
    eh0 = o._EH(0.31, 1e-9, 0.68, 0.965, 1.0)
    assert eh.transeh(0.3) < eh0.transeh(0.3)          # baryons suppress small scales
    # with no baryons T depends on k/h only through k/(Omega_m h^2)
    a = o._EH(0.25, 1e-12, 0.8, 1, 1)
    b = o._EH(0.4, 1e-12, 0.5, 1, 1)
    for kh in (0.01, 0.1, 1.0):          # same k/h in both
        assert a.transeh(kh * 0.8) == pytest.approx(b.transeh(kh * 0.5), rel=1e-6)


# Eisenstein & Hu 1998 (ApJ 496, 605) define the no-wiggle transfer function
# piece by piece; each test below pins one published piece of the oracle's
# composite transeh through a limit or an inversion, so that a 1% change of any
# constant in the oracle's formula fails at least one of them
# (tests/mutate_oracle_constants.py runs that mutation check).
E2 = 2 * math.e
THETA2 = (2.728 / 2.7) ** 2          # eq. 28: q = k Theta_2.7^2 / Gamma (k in 1/Mpc)


def _q_of(k, om, h, shape=None):
    return (k / h) * THETA2 / (shape if shape is not None else om * h)


def _T0(q):
    """EH98 eqs. 29: L0 / (L0 + C0 q^2), C0 = 14.2 + 731 / (1 + 62.5 q)."""
    L = math.log(E2 + 1.8 * q)
    return L / (L + (14.2 + 731 / (1 + 62.5 * q)) * q * q)


def _C_of(T, q):
    """C0(q) recovered from T0 = L0 / (L0 + C0 q^2)."""
    return (1 / T - 1) * math.log(E2 + 1.8 * q) / (q * q)


def _shape_from_T(T, k, h):
    """Invert T0 (monotone decreasing in q) by bisection: the effective shape
    Gamma_eff with T = T0(q(k, Gamma_eff))."""
    lo, hi = 1e-12, 1e12
    for _ in range(400):
        mid = math.sqrt(lo * hi)
        if _T0(mid) > T:
            lo = mid
        else:
            hi = mid
        if hi / lo < 1 + 1e-15:
            break
    q = math.sqrt(lo * hi)
    return (k / h) * THETA2 / q


def test_transeh_large_q_asymptote():
    """T0 -> L0 / (14.2 q^2) as q -> infinity (C0 -> 14.2), q from eq. 28 with
    Gamma = Omega_m h when there are no baryons."""
    om, h = float(np.float32(0.31)), float(np.float32(0.68))  # the reference's float parameters
    eh = o._EH(om, 0.0, h, 0.965, 1.0)
    for k in (1e5, 1e6, 1e7):
        q = _q_of(k, om, h)
        ratio = float(eh.transeh(k)) * 14.2 * q * q / math.log(E2 + 1.8 * q)
        assert ratio == pytest.approx(1.0, abs=3e-6), (k, ratio)


def test_transeh_small_q_limit():
    """C0(q -> 0) = 14.2 + 731 = 745.2, and L0(0) = ln(2e)."""
    om, h = float(np.float32(0.31)), float(np.float32(0.68))  # the reference's float parameters
    eh = o._EH(om, 0.0, h, 0.965, 1.0)
    k = 1e-6 * om * h * h / THETA2   # q = 1e-6
    q = _q_of(k, om, h)
    C = _C_of(float(eh.transeh(k)), q)
    assert C == pytest.approx(745.2 * (1 - 62.5e-6 * 731 / 745.2), rel=2e-5)


def test_transeh_crossover_of_C0():
    """C0 is halfway between its limits, 14.2 + 731/2, at q = 1/62.5."""
    om, h = float(np.float32(0.31)), float(np.float32(0.68))  # the reference's float parameters
    eh = o._EH(om, 0.0, h, 0.965, 1.0)
    q = 1 / 62.5
    k = q * om * h * h / THETA2
    assert _C_of(float(eh.transeh(k)), _q_of(k, om, h)) == pytest.approx(14.2 + 731 / 2, rel=1e-9)


def _alpha_gamma(om, ob, h):
    """EH98 eq. 31."""
    wm, fb = om * h * h, ob / om
    return 1 - 0.328 * math.log(431 * wm) * fb + 0.38 * math.log(22.3 * wm) * fb * fb


def _sound_horizon(om, ob, h):
    """EH98 eq. 26 (Mpc)."""
    wm, wb = om * h * h, ob * h * h
    return 44.5 * math.log(9.83 / wm) / math.sqrt(1 + 10 * wb ** 0.75)


@pytest.mark.parametrize("om,ob,h", [(0.31, 0.049, 0.68), (0.25, 0.06, 0.72), (0.4, 0.02, 0.6)])
def test_transeh_shape_limits_with_baryons(om, ob, h):
    """EH98 eq. 30: Gamma_eff = Omega_m h (alpha + (1 - alpha) / (1 + (0.43 k s)^4)):
    Omega_m h alpha at k s >> 1 (eq. 31), Omega_m h (1 + alpha)/2 where
    0.43 k s = 1 and Omega_m h (alpha + (1 - alpha)/17) where it is 2 (s from
    eq. 26).  The shapes are recovered by inverting T0."""
    eh = o._EH(om, ob, h, 0.965, 1.0)
    # float32 parameters, as the reference stores them
    om32, ob32, h32 = (float(np.float32(v)) for v in (om, ob, h))
    a = _alpha_gamma(om32, ob32, h32)
    s = _sound_horizon(om32, ob32, h32)
    base = om32 * h32
    k_big = 1e4 / (0.43 * s)
    assert _shape_from_T(float(eh.transeh(k_big)), k_big, h32) == pytest.approx(base * a, rel=1e-9)
    k_t = 1 / (0.43 * s)
    assert _shape_from_T(float(eh.transeh(k_t)), k_t, h32) == pytest.approx(base * (1 + a) / 2, rel=1e-9)
    k_2 = 2 / (0.43 * s)
    assert _shape_from_T(float(eh.transeh(k_2)), k_2, h32) == pytest.approx(base * (a + (1 - a) / 17), rel=1e-9)
    assert eh.transeh(1e-7) == pytest.approx(1.0, abs=1e-5)


def test_transeh_no_baryons_alpha_is_one():
    """Omega_b = 0: alpha = 1 and Gamma_eff = Omega_m h at every k."""
    om, h = 0.3, 0.7
    eh = o._EH(om, 0.0, h, 0.965, 1.0)
    for k in (1e-3, 0.05, 1.0, 30.0):
        om32, h32 = float(np.float32(om)), float(np.float32(h))
        assert _shape_from_T(float(eh.transeh(k)), k, h32) == pytest.approx(om32 * h32, rel=1e-9)


# Ma, Caldwell, Bode & Wang 1999 (ApJ 521, L1): g(Omega_m, w) = g(Omega_m, -1)
# (-w)^t, t = -(0.255 + 0.305 w + 0.0027/w)(1 - Omega_m)
#             - (0.366 + 0.266 w - 0.07/w) ln Omega_m;
# Carroll, Press & Turner 1992: g(Omega, Lambda) = 2.5 Omega / (Omega^(4/7) -
# Lambda + (1 + Omega/2)(1 + Lambda/70)).  At high z, Omega(z) -> 1 and
# t(z) -> 0, so the oracle's D(z)(1 + z) -> 1 / g0 with g0 = g(Omega_m, Omega_L)
# (-w)^t0: both fits are recovered from the oracle's D and compared with the
# published coefficients.
def _inv_g0(om, ol, w):
    z = 1e7  # Omega(z) = 1 - O(curvature / (Omega_m z)): float-exact only far out
    return float(o.growth(z, om, ol, w)) * float(np.float32(np.float32(1) + np.float32(z)))


def test_growth_cpt_coefficients_recovered():
    """w = -1: 1/g0 = (Omega^p - Lambda + (1 + Omega/2)(1 + Lambda/c)) / (2.5 Omega);
    recover p and 1/c by least squares over Omega_m, Omega_L (including open
    and closed models, which the fitting form allows)."""
    rows, rhs = [], []
    for om in (0.1, 0.2, 0.3, 0.5, 0.7, 0.9):
        for ol in (0.0, 0.3, 0.6, 0.9):
            om32, ol32 = float(np.float32(om)), float(np.float32(ol))
            y = _inv_g0(om, ol, -1.0) * 2.5 * om32 + ol32 - (1 + om32 / 2)
            # y = Omega^p + (1 + Omega/2) Lambda / c: linear in 1/c given p
            rows.append((om32, ol32, y))
    best = None
    for p in np.linspace(0.5, 0.65, 15001):
        A = np.array([[(1 + r[0] / 2) * r[1]] for r in rows])
        b = np.array([r[2] - r[0] ** p for r in rows])
        inv_c, res, *_ = np.linalg.lstsq(A, b, rcond=None)
        err = float(np.sum((A @ inv_c - b) ** 2))
        if best is None or err < best[0]:
            best = (err, p, float(inv_c[0]))
    _, p, inv_c = best
    assert p == pytest.approx(4 / 7, abs=2e-4)
    assert inv_c == pytest.approx(1 / 70, rel=3e-3)


def test_growth_ma_w_coefficients_recovered():
    """t0 = ln(g0 / g_cpt) / ln(-w) over a grid of (Omega_m, w); the six
    coefficients of the Ma et al. fit by least squares."""
    A, b = [], []
    for om in (0.1, 0.2, 0.3, 0.45, 0.6, 0.8):
        for w in (-0.3, -0.4, -0.5, -0.6, -0.7, -1.4, -1.8, -2.5, -3.5):
            om32, ol32, w32 = float(np.float32(om)), float(np.float32(1 - om)), float(np.float32(w))
            g_cpt = 2.5 * om32 / (om32 ** (4 / 7) - ol32 + (1 + om32 / 2) * (1 + ol32 / 70))
            g0 = 1 / _inv_g0(om, 1 - om, w)
            t0 = math.log(g0 / g_cpt) / math.log(-w32)
            lo = math.log(float(np.float32(om32)))
            A.append([-(1 - om32), -(1 - om32) * w32, -(1 - om32) / w32, -lo, -lo * w32, -lo / w32])
            b.append(t0)
    coef, *_ = np.linalg.lstsq(np.array(A), np.array(b), rcond=None)
    pub = np.array([0.255, 0.305, 0.0027, 0.366, 0.266, -0.07])
    assert np.allclose(coef, pub, rtol=2e-3, atol=2e-5), coef


def test_growth_w_fit_tracks_the_growth_equation_in_matter_normalisation():
    """1/g0 is D(z -> inf)(1+z): compare the growth suppression g0 of the fit with
    that of the wCDM growth equation (D = a early on) at w != -1 (a fitting
    formula: agreement to 1.5%)."""
    for om, w in ((0.31, -0.8), (0.25, -0.6), (0.35, -1.3)):
        exact = _growth_exact(0.0, om, 1 - om, w) / _growth_exact(2000.0, om, 1 - om, w) / 2001.0
        assert 1 / _inv_g0(om, 1 - om, w) == pytest.approx(exact, rel=0.015)


def test_deltac_einstein_de_sitter():
    """Omega_m = 1: delta_c = 3 (12 pi)^(2/3) / 20 = 1.686 times (1 + z)."""
    dc0 = 3 * (12 * math.pi) ** (2 / 3) / 20
    for z in (0.0, 6.0, 20.0):
        assert float(o.deltac(z, 1.0, 0.0, -1.0)) == pytest.approx(dc0 * (1 + z), rel=3e-7)


def test_deltac_open_branch_tends_to_eds():
    """The open-universe form (Omega_L = 0, Omega_m < 1) tends to the EdS value as
    Omega(z) -> 1 (high z); at z = 0 with Omega_m = 0.3 it is below 1.686."""
    dc0 = 3 * (12 * math.pi) ** (2 / 3) / 20
    # Omega(z = 50) = 0.956, where the O(eta^2) terms are ~1e-3; further out the
    # reference's float sinh(eta) - eta cancels catastrophically
    hi = float(o.deltac(50.0, 0.3, 0.0, -1.0)) * float(o.growth(50.0, 0.3, 0.0, -1.0))
    assert hi == pytest.approx(dc0, rel=2e-3)
    lo = float(o.deltac(0.0, 0.3, 0.0, -1.0))
    assert 1.6 < lo < dc0


def test_deltac_flat_omega_dependence_recovered():
    """Flat models (Kitayama & Suto 1996): delta_c(Omega) = 1.686 (1 + 0.0123 log10
    Omega); the slope is recovered from delta_c(z = 0) D(0) over Omega_m."""
    dc0 = 3 * (12 * math.pi) ** (2 / 3) / 20
    xs, ys = [], []
    for om in (0.1, 0.2, 0.3, 0.5, 0.7, 0.9):
        om32 = float(np.float32(om))
        ol32 = float(np.float32(np.float32(1) - np.float32(om)))
        xs.append(math.log10(om32))
        ys.append(float(o.deltac(0.0, om32, ol32, -1.0)) / dc0 - 1)
    slope = float(np.polyfit(xs, ys, 1)[0])
    assert slope == pytest.approx(0.0123, rel=1e-3)


def test_fcoll_is_erfc_of_the_threshold():
    """fcoll = erfc(delta_c(z) / (sqrt 2 sigma_min)): 1 at sigma_min -> inf,
    -> 0 at large z, and the closed form at one point (EdS)."""
    dc0 = 3 * (12 * math.pi) ** (2 / 3) / 20
    v = float(o.fcoll(3.0, 1.0, 0.0, -1.0, 5.0))
    assert v == pytest.approx(math.erfc(dc0 * 4 / math.sqrt(2) / 5.0), rel=2e-6)
    assert float(o.fcoll(40.0, 0.31, 0.69, -1.0, 5.6)) < 1e-6


def test_sigma_integral_in_k_matches_ln_k():
    """sigma^2 = pnorm^2 int k^(2+ns) T^2 W^2 dk (the ln k Jacobian)."""
    eh = o._EH(0.31, 0.049, 0.68, 0.965, 1.0)
    R = 5.0
    ln_form = eh.integral(R)

    def f(k):
        return k ** (2 + 0.965) * eh.transeh(k) ** 2 * o.tophat_W(np.array([k * R]))[0] ** 2

    pts = np.logspace(-5, 5, 41) / R
    k_form = sum(integrate.quad(f, a, b, epsrel=1e-10, limit=400)[0] for a, b in zip(pts[:-1], pts[1:]))
    assert ln_form == pytest.approx(k_form, rel=1e-6)


def test_sigma8_identity():
    pn = o.normalize(P.Omegam, P.Omegab, P.h, P.ns, P.sigma8)
    R8 = 8.0 / P.h
    M8 = 4 * math.pi / 3 * 2.775e11 * P.Omegam * P.h ** 2 * R8 ** 3
    s8 = math.sqrt(o.cdmehint(P.Omegam, P.Omegab, P.h, P.ns, pn, M8))
    assert s8 == pytest.approx(P.sigma8, rel=2e-6)


def test_sigma_decreases_with_radius_and_stays_below_sigma_min():
    sc = o.setup_scalars(o.D2ZParams(N=64, BoxSize=100.0, Nscales=10, Rmax=25.0))
    assert np.all(np.diff(sc.sigma) < 0)
    assert np.all(sc.sigma < sc.sigma_min)


# -------------------------------------------------------------- erfc^-1
@pytest.mark.parametrize("zeta", [10.0, 40.0, 100.0, 3.0])
def test_inverse_erfc_is_the_lower_end_of_a_tight_bracket(zeta):
    x = np.float32(1.0 / np.float32(zeta))
    root = special.erfcinv(float(x))
    e = float(o.inverse_erfc(1.0 / zeta))
    assert root - 1.1e-4 * root <= e <= root + 1e-6
    assert math.erfc(float(np.float32(e))) >= float(x) * (1 - 1e-6)
    assert root == pytest.approx(optimize.brentq(lambda t: math.erfc(t) - float(x), 0, 100, xtol=1e-15),
                                 abs=1e-12)


# ---------------------------------------------------------------- barrier
def test_barrier_solves_the_collapse_condition():
    """z from the affine closed form satisfies zeta f_coll(delta_R, z) = 1 with
    f_coll = erfc((deltac/D(z) - delta_R)/sqrt(2(smin^2 - s^2))) and matter-era
    D(z) = D(20) 21/(1+z)."""
    p = o.D2ZParams(N=64, BoxSize=100.0, Nscales=6, Rmax=25.0, zeta=40.0)
    sc = o.setup_scalars(p)
    rng = np.random.default_rng(2)
    for r in range(p.Nscales):
        dR = rng.normal(0, 2, 20)
        z = o.barrier(dR, sc.A, sc.B[r])
        D = float(sc.DHigh) * 21 / (1 + z)
        s = math.sqrt(2 * (float(sc.sigma_min) ** 2 - float(sc.sigma[r]) ** 2))
        lhs = p.zeta * special.erfc((1.686 / D - dR) / s)
        np.testing.assert_allclose(lhs, 1.0, rtol=2e-3)


def test_barrier_known_answer_at_zero_density():
    sc = o.setup_scalars(o.D2ZParams(N=32, BoxSize=50.0, Nscales=3, Rmax=10.0))
    z0 = o.barrier(np.zeros(1), sc.A, sc.B[0])[0]
    assert z0 == pytest.approx(float(sc.A) * sc.B[0] - 1, rel=1e-15)


# -------------------------------------------------------- running maximum
def test_running_max_is_zero_floored_max_over_radii():
    N = 16
    p = o.D2ZParams(N=N, BoxSize=25.0, Nscales=5, Rmax=6.0, zeta=40.0)
    sc = o.setup_scalars(p)
    d = grf.grf(N, 25.0, 3)
    z, _, _ = o.delta2zreion(d, p, sc, return_fp64=True)
    dhat = o.forward(d, sc.DInit)
    k = o.kmag_grid(N, sc.box_mpc)
    zs = [o.barrier(o.smooth(dhat, k, sc.R[r]), sc.A, sc.B[r]) for r in range(p.Nscales)]
    np.testing.assert_array_equal(z, np.maximum(0.0, np.max(zs, axis=0)))


# ---------------------------------------------------------------- history
def test_history_counts_brute_force():
    rng = np.random.default_rng(4)
    z = np.concatenate([rng.uniform(-1, 55, 3000), [0.0, 49.9, 49.9004, 0.0105, 0.009]]).astype(np.float32)
    dz = float(o.history_dz())
    ref = np.zeros(o.NHISTORY, np.int64)
    for v in z:
        q = (float(v) - 49.9) / dz
        b = int(q)  # C truncation toward zero
        if 0 <= b < o.NHISTORY:
            ref[b] += 1
    np.testing.assert_array_equal(o.history_counts(z), ref)


def test_tau_of_instant_reionization_matches_quadrature():
    zr = 8.0
    N = 8
    z = np.full(N ** 3, zr, np.float32)
    counts = o.history_counts(z)
    x = o.history_float(counts, N)
    za = o.history_z()
    t = float(o.tau(za, x, 0.31, 0.049, 0.68))
    h = 0.68
    ne0 = 0.049 * h * h * 1.88e-29 / 1.67e-24
    dtau0 = 6.65e-25 * ne0 * 3e10 / (100.0 / 3.086e19 * h)

    def f(zz):
        nhe = 2 if zz < 3 else 1
        return (1 + zz) ** 2 / math.sqrt(0.31 * (1 + zz) ** 3 + 1 - 0.31) * (1 - 0.25 + nhe * 0.25 / 4)

    ref = dtau0 * (integrate.quad(f, 0.01, 3)[0] + integrate.quad(f, 3, zr)[0])
    assert t == pytest.approx(ref, rel=2e-3)


# --------------------------------------------------------------------- GRF
def test_philox_known_answer_vectors():
    # Random123 known-answer vectors for philox4x32-10
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, want in cases:
        got = grf.philox4x32_10(*[np.array([c]) for c in ctr], *key)
        assert tuple(int(g[0]) for g in got) == want


def test_normals_are_standard():
    g = grf.cell_normals(7, 0, 200000).astype(np.float64)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01
    assert abs(np.mean(g ** 4) - 3) < 0.06


def test_grf_spectrum_and_normalisation():
    # sigma(8 Mpc/h) of the spectrum from an independent k-space quadrature
    A = grf.amplitude()
    gam = grf.bbks_gamma(0.31, 0.049, 0.68)

    def f(k):
        return A * k ** 2 * k ** 0.965 * grf.t_bbks(k, gam) ** 2 * grf._w_tophat(8 * k) ** 2 / (2 * np.pi ** 2)

    pts = np.logspace(-6, 4, 51)
    s2 = sum(integrate.quad(f, a, b, limit=200)[0] for a, b in zip(pts[:-1], pts[1:]))
    assert math.sqrt(s2) == pytest.approx(0.81, rel=1e-6)
    assert float(grf.t_bbks(1e-8, gam)) == pytest.approx(1.0, abs=1e-6)
    # measured power of a 64^3 field against P(k) in shells
    N, box = 64, 100.0
    d = grf.grf(N, box, 11).astype(np.float64)
    pk = np.abs(np.fft.rfftn(d)) ** 2 * (box / N) ** 3 / N ** 3
    kk = np.fft.fftfreq(N, 1.0 / N)
    kz = np.arange(N // 2 + 1)
    kmag = np.sqrt(kk[:, None, None] ** 2 + kk[None, :, None] ** 2 + kz[None, None, :] ** 2) * 2 * np.pi / box
    for lo, hi in [(0.2, 0.4), (0.4, 0.8), (0.8, 1.6)]:
        sel = (kmag >= lo) & (kmag < hi)
        ratio = np.mean(pk[sel]) / np.mean(grf.power(kmag[sel]))
        assert ratio == pytest.approx(1.0, rel=0.08)


# --------------------------------------------------- two restatements agree
@pytest.mark.parametrize("N", [16, 32])
def test_c_restatement_agrees_with_fp64_oracle(N):
    box = 1.5625 * N
    p = o.D2ZParams(N=N, BoxSize=box, Nscales=6, Rmax=box / 4, zeta=40.0)
    sc = o.setup_scalars(p)
    d = grf.grf(N, box, 1)
    zc, sgc = d2z_cpu.run(d, p, sc, nthreads=2)
    zr, sgr, _ = o.delta2zreion(d, p, sc)
    assert np.max(np.abs(zc.astype(np.float64) - zr)) < 1e-4
    np.testing.assert_allclose(sgc, sgr, rtol=1e-5)


def test_radius_schedule_endpoints_and_spacing():
    p = o.D2ZParams(N=128, BoxSize=200.0, Nscales=32, Rmax=50.0)
    R = o.radius_schedule(p).astype(np.float64)
    assert R[0] == pytest.approx(200.0 / 0.68 / 128, rel=1e-6)
    assert R[-1] == pytest.approx(50.0 / 0.68, rel=1e-5)
    np.testing.assert_allclose(np.diff(np.log(R)), np.log(R[-1] / R[0]) / 31, rtol=1e-4)


# ---------------------------------------------------- displacement (Delta2Velocity)
def _plane(N, n):
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return 2.0 * np.pi * (n[0] * i + n[1] * j + n[2] * k) / N


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_displacement_of_a_plane_wave_is_closed_form(axis):
    """delta = cos(k.x) gives psi_a = -(k_a / k^2) sin(k.x): pins the sign of i,
    the 1/N^3 normalisation, the axis order and k = 2 pi n / box."""
    N, box = 16, 37.0
    n = (2, -3, 5)
    ph = _plane(N, n)
    dk = 2.0 * np.pi / box
    k = np.array(n, dtype=np.float64) * dk
    psi = o.displacement(np.cos(ph), box, axis)
    ref = -(k[axis] / np.dot(k, k)) * np.sin(ph)
    assert np.max(np.abs(psi - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_displacement_zeroes_the_components_own_nyquist_plane():
    """x Nyquist mixed with a y mode: the x component is zeroed (the reference
    skips kx = N/2 for xdv_x), the y component is not."""
    N, box = 16, 20.0
    ph = _plane(N, (N // 2, 1, 0))
    delta = np.cos(ph)
    dk = 2.0 * np.pi / box
    kx, ky = dk * N / 2, dk * 1.0
    assert np.max(np.abs(o.displacement(delta, box, 0))) <= 1e-14
    ref_y = -(ky / (kx * kx + ky * ky)) * np.sin(ph)
    assert np.max(np.abs(o.displacement(delta, box, 1) - ref_y)) <= 1e-12 * np.max(np.abs(ref_y))
    assert np.max(np.abs(o.displacement(delta, box, 2))) <= 1e-14


# ------------------------------------------------------------ PowerSpectrum
def test_power_spectrum_of_a_plane_wave_is_closed_form():
    """cos(k0.x) puts |delta_k|^2 = (N^3/2 N^-1.5)^2 = N^3/4 into the one mode
    +k0 of the half spectrum: ps of its bin = (2 pi)^3 (N^3/4) / d3k / count."""
    N, box = 16, 30.0
    n0 = (3, 1, 2)
    delta = np.cos(_plane(N, n0))
    dk = 2.0 * np.pi / box
    k0 = dk * np.sqrt(np.dot(n0, n0))
    nk, kmin, kmax = 12, 0.93 * dk, 0.47 * dk * N * np.sqrt(3.0)
    r = o.power_spectrum(delta, box, nk, kmin, kmax)
    b0 = int((np.log(k0) - np.log(kmin)) / ((np.log(kmax) - np.log(kmin)) / nk))
    d3k = np.float64(np.float32(np.float64(np.float32(np.float32(dk) * N)) ** 3))
    assert r["ps"][b0] == pytest.approx((2 * np.pi) ** 3 * (N ** 3 / 4.0) / d3k / r["nmodes"][b0], rel=1e-12)
    others = [b for b in range(nk) if b != b0 and r["nmodes"][b] > 0]
    assert max(r["ps"][others]) <= 1e-20 * r["ps"][b0]


def test_power_spectrum_mode_counts_brute_force():
    """Counts of the half-spectrum modes per log bin against an independent
    FP64 lattice count (the bins chosen so no mode sits within 1e-4 of an
    edge), including the reference's truncation of u in (-1, 0) into bin 0."""
    N, box, nk = 16, 25.0, 9
    dk = 2.0 * np.pi / box
    kmin, kmax = 1.37 * dk, 6.91 * dk
    r = o.power_spectrum(np.random.default_rng(0).standard_normal((N, N, N)), box, nk, kmin, kmax)
    dl = (np.log(kmax) - np.log(kmin)) / nk
    counts = np.zeros(nk, dtype=np.int64)
    for nx in range(-N // 2, N // 2):
        for ny in range(-N // 2, N // 2):
            for nz in range(0, N // 2 + 1):
                nzs = -N // 2 if nz == N // 2 else nz
                k = dk * np.sqrt(nx * nx + ny * ny + nzs * nzs)
                if k == 0:
                    continue
                u = (np.log(k) - np.log(kmin)) / dl
                assert abs(u - round(u)) > 1e-4
                if -1.0 < u < nk:
                    counts[int(u)] += 1
    assert np.array_equal(r["nmodes"], counts)
