"""GPU parity of the binned power spectrum (PowerSpectrum,
c++/src/powerspectrum_lmb.C:8-139) through the C ABI, against the oracle
(float binning as the reference, FP64 power).

Tolerance: mode counts exact (the bins are chosen so that no mode lies within
1e-5 of a bin edge, where a one-ulp difference of the float logarithm could
move it); ps, <1/k^2> and <k> within 1e-5 relative; empty bins NaN in both.
"""
import numpy as np
import pytest

import d2z_fp64 as oracle
from _cases import field, lib_params
from conftest import announce
from prc_d2z import D2Z, D2ZError, FUSED_FP32, FUSED_MIXED, PLAIN_FP32

pytestmark = pytest.mark.gpu


def _bins(N, box, nk):
    """kmin, kmax with every binned mode at least 1e-5 from a bin edge."""
    dk = 2.0 * np.pi / box
    kmin, kmax = 1.07 * dk, 0.93 * dk * N / 2 * np.sqrt(3.0)
    probe = np.zeros((N, N, N))
    for _ in range(50):
        u = oracle.power_spectrum(probe, box, nk, kmin, kmax)["u"]
        v = u[(u > -1.0) & (u < nk)].astype(np.float64)
        if np.min(np.abs(v - np.round(v))) > 1e-5:
            return kmin, kmax
        kmin *= 1.0003
    raise AssertionError("no bin set clear of the modes")


@pytest.mark.parametrize("N,mode", [(32, FUSED_FP32), (64, PLAIN_FP32), (128, FUSED_MIXED), (192, FUSED_FP32)])
def test_power_spectrum_vs_oracle(N, mode):
    lp = lib_params(N, mode=mode)
    nk = 20
    kmin, kmax = _bins(N, lp.BoxSize, nk)
    delta = field(N)
    announce(f"power spectrum N={N} mode={mode}")
    with D2Z(lp) as d:
        d.set_delta(delta)
        got = d.power_spectrum(nk, kmin, kmax)
    ref = oracle.power_spectrum(delta, lp.BoxSize, nk, kmin, kmax)
    assert np.array_equal(got["nmodes"], ref["nmodes"])
    for key in ("ps", "ook2", "kmean"):
        np.testing.assert_allclose(got[key], ref[key], rtol=1e-5, equal_nan=True, err_msg=key)
    print(f"N={N}: {int(ref['nmodes'].sum())} modes binned, max rel ps diff "
          f"{np.nanmax(np.abs(got['ps'] / ref['ps'] - 1)):.2e}")


def test_power_spectrum_plane_wave_and_errors():
    N = 64
    lp = lib_params(N)
    n0 = (5, -3, 7)
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    delta = np.cos(2.0 * np.pi * (n0[0] * i + n0[1] * j + n0[2] * k) / N).astype(np.float32)
    nk = 16
    kmin, kmax = _bins(N, lp.BoxSize, nk)
    dk = 2.0 * np.pi / lp.BoxSize
    b0 = int((np.log(dk * np.sqrt(np.dot(n0, n0))) - np.log(kmin)) / ((np.log(kmax) - np.log(kmin)) / nk))
    announce("power spectrum plane wave N=64")
    with D2Z(lp) as d:
        d.set_delta(delta)
        got = d.power_spectrum(nk, kmin, kmax)
        for bad in ((0, kmin, kmax), (513, kmin, kmax), (nk, kmax, kmin), (nk, 0.0, kmax)):
            with pytest.raises(D2ZError) as ei:
                d.power_spectrum(*bad)
            assert ei.value.code == -1
    d3k = np.float64(np.float32(np.float64(np.float32(np.float32(dk) * N)) ** 3))
    assert got["ps"][b0] == pytest.approx((2 * np.pi) ** 3 * (N ** 3 / 4.0) / d3k / got["nmodes"][b0], rel=1e-5)
    others = [b for b in range(nk) if b != b0 and got["nmodes"][b] > 0]
    assert np.max(got["ps"][others]) <= 1e-10 * got["ps"][b0]
