"""Pins of the history-correction oracle (oracle/bubble_cpu.c via oracle/bubble.py).

Every check compares the oracle with something it does not compute itself:
closed forms (linear tables, a constant-rate recurrence, the identity
correction), and an exact enumeration of the skewer statistics -- every
admissible start cell, its walk lengths along +y and +z found by whole-array
shifts, and the mean and mass of the log-uniform distance per walk length --
against which the oracle's Monte Carlo must agree within its sampling error.
"""
import math

import numpy as np
import pytest

import bubble as ob

DZH = np.float32((0.01 - 49.9) / (ob.NHISTORY - 1))
DZC = np.float32((0.01 - 49.9) / (ob.NCHISTORY - 1))


def test_table_value_linear_and_edges():
    n, tmin, tmax = 64, 2.0, 34.0
    delta = (tmax - tmin) / n
    t = (3.0 - 0.25 * (tmin + delta * np.arange(n))).astype(np.float32)
    for v in np.linspace(tmin + 0.01, tmin + (n - 1) * delta - 0.01, 37):
        assert ob.table_value(t, tmin, tmax, v) == pytest.approx(3.0 - 0.25 * v, abs=2e-6)
    assert ob.table_value(t, tmin, tmax, tmin - 5) == t[0]
    assert ob.table_value(t, tmin, tmax, tmin) == t[0]
    assert ob.table_value(t, tmin, tmax, tmax + 1) == t[-1]
    # the last interval holds the last value (the reference reads past the table)
    assert ob.table_value(t, tmin, tmax, tmax - 0.1) == t[-1]


def test_mfp_interpolation_linear_and_edges():
    x = (0.01 + np.arange(ob.NX_MFP) * np.float32(0.989 / 999)).astype(np.float32)
    mfp = (5.0 + 20.0 * x.astype(np.float64)).astype(np.float32)
    for v in (0.0123, 0.2, 0.5, 0.77, 0.99):
        assert ob.mfp_at(mfp, v) == pytest.approx(5.0 + 20.0 * v, rel=2e-6)
    assert ob.mfp_at(mfp, 0.001) == mfp[0]
    assert ob.mfp_at(mfp, 0.9995) == mfp[-1]
    assert ob.mfp_at(np.full(ob.NX_MFP, 7.5, np.float32), 0.42) == np.float32(7.5)


def test_chistory_constant_rate_closed_form():
    # fcoll(z) = a (50 - z) and a constant mfp: every step adds
    # a |dzc| zeta / (1 + mfp/rmfp) until the clamp at 1
    a, fmin, fmax, n = 1e-3, 0.0, 50.0, 5000
    zt = fmin + (fmax - fmin) / n * np.arange(n)
    t = (a * (50.0 - zt)).astype(np.float32)
    mfp = np.full(ob.NX_MFP, 12.0, np.float32)
    rmfp, zeta = 30.0, 40.0
    ch = ob.corrected_history(t, fmin, fmax, mfp, rmfp, zeta)
    step = a * abs(float(DZC)) * zeta / (1.0 + 12.0 / rmfp)
    i = np.arange(ob.NCHISTORY)
    expect = np.minimum(i * step, 1.0)
    assert ch[0] == 0.0
    assert np.max(np.abs(ch - expect)) < 2e-3
    assert np.all(np.diff(ch) >= 0) and ch.max() == np.float32(1.0)
    # a rate that depends on the current fraction: larger bubbles slow the growth
    mfp2 = (12.0 + 100.0 * (0.01 + np.arange(ob.NX_MFP) * 0.989 / 999)).astype(np.float32)
    ch2 = ob.corrected_history(t, fmin, fmax, mfp2, rmfp, zeta)
    assert np.all(ch2 <= ch + 1e-6) and ch2[-1] < ch[-1]


def test_correction_identity_when_histories_agree():
    # chistory sampled from the same x(z) as the history: the corrected
    # redshift of every cell is its own to within one history bin
    zh = 49.9 + np.arange(ob.NHISTORY) * np.float64(DZH)
    zc = 49.9 + np.arange(ob.NCHISTORY) * np.float64(DZC)
    X = lambda z: (1.0 - np.exp(-(49.9 - z) / 8.0)) / (1.0 - math.exp(-49.89 / 8.0))
    hist = X(zh).astype(np.float32)
    ch = X(zc).astype(np.float32)
    zvz = ob.zvz_table(hist, ch)
    ok = hist > 1e-3
    assert np.max(np.abs(zvz[ok] - zh[ok])) < 2e-3
    rng = np.random.default_rng(3)
    z = rng.uniform(1.0, 40.0, size=(16, 16, 16)).astype(np.float32)
    out = ob.correct_zreion(z, hist, ch)
    assert np.max(np.abs(out - z)) < 2e-3
    # linear histories: the correction is the identity up to the bin width
    lin_h = (np.arange(ob.NHISTORY) / (ob.NHISTORY - 1)).astype(np.float32)
    lin_c = (np.arange(ob.NCHISTORY) / (ob.NCHISTORY - 1)).astype(np.float32)
    out = ob.correct_zreion(z, lin_h, lin_c)
    assert np.max(np.abs(out - z)) <= 1.01 * abs(float(DZH)) + 1e-4


def test_zvz_fill_and_clamps():
    hist = np.zeros(ob.NHISTORY, np.float32)
    hist[500:] = np.linspace(0.05, 0.95, ob.NHISTORY - 500).astype(np.float32)
    ch = np.linspace(0.1, 0.9, ob.NCHISTORY).astype(np.float32)
    raw = ob.zvz_table(hist, ch, filled=False)
    assert np.all(raw[:500] == -1.0)
    assert raw[500] == np.float32(49.9)        # below chistory[0]
    assert raw[-1] == np.float32(0.01)         # above chistory[-1]
    filled = ob.zvz_table(hist, ch)
    assert np.all(filled[:500] == raw[500])
    # out-of-table cells take the first and last bins
    z = np.array([60.0, 49.95, 0.005, 0.0, 25.0], np.float32).reshape(5, 1, 1)
    out = ob.correct_zreion(np.broadcast_to(z, (5, 1, 4)), hist, ch)
    assert np.all(out[0] == filled[0]) and np.all(out[1] == filled[0])
    assert np.all(out[2] == filled[-1]) and np.all(out[3] == filled[-1])


def _field(N, seed=5):
    """Continuous zreion with different structure along y (sawtooth) and z."""
    rng = np.random.default_rng(seed)
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    P = 8
    z = (10.0 + 2.0 * ((P - 1 - (j % P)) / P) + 0.6 * np.sin(2 * np.pi * (k + 2 * i) / N)
         + 1e-3 * rng.standard_normal((N, N, N)))
    return z.astype(np.float32)


def _expected(z, zc):
    """Exact mean distance (cells) and P(m >= 2) over admissible starts, both walks."""
    N = z.shape[0]
    i, j, k = np.nonzero(z > zc)
    keep = (j < N - 2) & (k < N - 2)
    i, j, k = i[keep], j[keep], k[keep]
    ms = []
    for axis in (1, 2):
        m = np.full(i.size, N)
        done = np.zeros(i.size, bool)
        for t in range(1, N + 1):
            jj = (j + t) % N if axis == 1 else j
            kk = (k + t) % N if axis == 2 else k
            stop = ~done & (z[i, jj, kk] < zc)
            m[stop] = t
            done |= stop
        ms.append(m)
    m = np.concatenate(ms).astype(np.float64)
    mean_d = np.where(m > 1, 1.0 / np.log(np.maximum(m, 2) / np.maximum(m - 1, 1)), 0.5 / math.log(2.0))
    # second moment of a log-uniform variable on [lo, hi]: (hi^2 - lo^2) / (2 ln(hi/lo))
    lo = np.where(m > 1, m - 1, 0.5)
    second = (m ** 2 - lo ** 2) / (2 * np.log(m / lo))
    return mean_d.mean(), second.mean() - mean_d.mean() ** 2, (m >= 2).mean(), i.size / z.size


def test_skewers_against_exact_enumeration():
    N, box, nsk = 32, 64.0, 4000
    z = _field(N)
    counts = np.bincount(np.clip(((z.astype(np.float64).ravel() - 49.9) / float(DZH)).astype(np.int64), 0,
                                 ob.NHISTORY - 1), minlength=ob.NHISTORY)
    hist = np.cumsum((counts / float(N) ** 3).astype(np.float32), dtype=np.float32)
    mfp, mfp_x, dfd, zcur, _ = ob.bubble_mfp(z, hist, box, nskewers=nsk, seed=11)
    cell = box / N
    assert mfp_x[0] == np.float32(0.01) and mfp_x[-1] == pytest.approx(0.999, abs=1e-6)
    for ix in (30, 200, 450, 700, 900):
        zc = float(zcur[ix])
        mean_d, var_d, p2, frac = _expected(z, zc)
        # zcur: the fraction of cells above it is the ionized fraction
        assert abs(float((z > zc).mean()) - float(mfp_x[ix])) < 3e-3
        sigma = math.sqrt(var_d / (2 * nsk))
        assert abs(mfp[ix] / cell - mean_d) < 5 * sigma + 1e-6, (ix, mfp[ix] / cell, mean_d)
        # walks of one cell fall below ln(cell) and are not binned
        n2 = 2 * nsk
        got = dfd[ix].sum()
        assert abs(got - n2 * p2) < 5 * math.sqrt(n2 * p2 * (1 - p2)) + 1, (ix, got, n2 * p2)
    # +y walks see the sawtooth: reversing the field along y changes the answer
    mfp_r, *_ = ob.bubble_mfp(z[:, ::-1, :], hist, box, nskewers=nsk, seed=11)
    assert abs(float(mfp_r[450]) - float(mfp[450])) > 0.05 * cell


def test_skewer_bins_of_whole_wraps():
    # uniform field: every walk wraps (m = N), distances in [N-1, N) cells,
    # binned just below ln BoxSize
    N, box = 16, 32.0
    rng = np.random.default_rng(2)
    z = (20.0 + 0.01 * rng.random((N, N, N))).astype(np.float32)
    b = int((20.0 - 49.9) / float(DZH))  # history bin of z = 20; the crossings sit below z = 20
    hist = np.zeros(ob.NHISTORY, np.float32)
    hist[b + 10:] = 0.5
    hist[b + 40:] = 1.0
    mfp, _, dfd, zcur, _ = ob.bubble_mfp(z, hist, box, nskewers=100, seed=3)
    assert np.all(zcur < 20.0)
    cell = box / N
    expect = cell / math.log(N / (N - 1.0))
    assert np.allclose(mfp, expect, rtol=0.02)
    lmin, lmax = math.log(cell), math.log(box)
    dm = (lmax - lmin) / ob.NMFP
    first = int((math.log((N - 1) * cell) - lmin) / dm)
    assert dfd[:, :first].sum() == 0 and np.all(dfd.sum(axis=1) == 200)


def test_skewers_fail_without_start_cells():
    z = np.full((8, 8, 8), 5.0, np.float32)
    hist = np.ones(ob.NHISTORY, np.float32)  # zcur = 49.9 everywhere: nothing above it
    with pytest.raises(RuntimeError):
        ob.bubble_mfp(z, hist, 16.0, nskewers=4, seed=1)
