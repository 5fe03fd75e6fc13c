"""GPU parity of the zreion history correction (main_d2z.C:78-90:
GetBubbleMFP, GetCHistory, CorrectHistory of c++/src/bubblehistory_d2z.C:44-343)
through the C ABI, against the oracle (oracle/bubble_cpu.c).

Both draw the same Philox4x32-10 numbers, so every skewer takes the same start
cell and the same walk lengths.  Tolerances: distances and their logarithms are
FP64 on both sides but from different libm implementations (one-ulp
differences), so the walk counts per ln-distance bin may differ where a
distance sits within ~1e-16 of a bin edge (at most 4 walks in all allowed) and
the mean free paths agree to 1e-6 relative; the corrected history (host float
arithmetic) and the corrected zreion (per-bin table and remap) are bit-exact.
"""
import numpy as np
import pytest

import bubble as ob
import d2z_fp64 as oracle
from _cases import box_mpc, field, lib_params
from conftest import announce
from prc_d2z import D2Z, D2ZError, D2ZParams, FUSED_FP32, FUSED_MIXED, NCHISTORY, NMFP, NX_MFP

pytestmark = pytest.mark.gpu


def _fcoll_table():
    """A collapsed-fraction table fcoll(z) on [0, 50]: smooth and decreasing in z."""
    n, fmin, fmax = 2000, 0.0, 50.0
    z = fmin + (fmax - fmin) / n * np.arange(n)
    return (0.03 * np.exp(-((z / 9.0) ** 2))).astype(np.float32), fmin, fmax


def _run(N, mode):
    lp = lib_params(N, nscales=16, mode=mode)
    d = D2Z(lp)
    d.set_delta(field(N))
    d.run()
    return lp, d


@pytest.mark.parametrize("N,mode,nsk", [(64, FUSED_FP32, 20000), (128, FUSED_MIXED, 4000)])
def test_history_correction_vs_oracle(N, mode, nsk):
    announce(f"history correction N={N} mode={mode} nskewers={nsk}")
    lp, d = _run(N, mode)
    with d:
        z = d.zreion()
        hist = D2Z.cumulative_history(d.history(), N)
        mfp, mfp_x, dfd = d.bubble_mfp(hist, nskewers=nsk, seed=7)
        # the walks measure Mpc: the reference divides BoxSize by h first (main_d2z.C:29)
        mfp_o, x_o, dfd_o, _, _ = ob.bubble_mfp(z, hist, box_mpc(lp), nskewers=nsk, seed=7)
        mfp_h, *_ = ob.bubble_mfp(z, hist, lp.BoxSize, nskewers=nsk, seed=7)
        assert not np.allclose(mfp, mfp_h, rtol=1e-3)  # Mpc/h units would differ by 1/h
        np.testing.assert_allclose(mfp_h / np.float32(lp.h), mfp_o, rtol=1e-5)  # x Mpc/h = x/h Mpc
        assert np.array_equal(mfp_x, x_o)
        assert dfd.shape == (NX_MFP, NMFP)
        assert np.abs(dfd.astype(np.int64) - dfd_o.astype(np.int64)).sum() <= 4
        assert dfd.sum() > 0
        np.testing.assert_allclose(mfp, mfp_o, rtol=1e-6)
        t, fmin, fmax = _fcoll_table()
        ch = d.corrected_history(t, fmin, fmax, mfp)
        ch_o = ob.corrected_history(t, fmin, fmax, mfp, lp.Rmax, lp.zeta)
        assert ch.shape == (NCHISTORY,)
        assert np.array_equal(ch, ch_o)
        d.correct_zreion(hist, ch)
        zc = d.zreion()
        zc_o = ob.correct_zreion(z, hist, ch)
        assert np.array_equal(zc, zc_o)
        # GetHistory of the corrected field, whose bins span more than one LDS
        # window of the histogram kernel
        b = oracle.history_bins(zc)
        assert b.max() - b.min() > 36000
        assert np.array_equal(d.history().astype(np.int64), oracle.history_counts(zc))
        # the correction maps each history bin to one redshift and keeps the order of cells
        o = np.argsort(z.ravel(), kind="stable")
        assert np.all(np.diff(zc.ravel()[o]) <= 0) or np.all(np.diff(zc.ravel()[o]) >= 0)
        print(f"N={N}: mfp(x=0.5) {mfp[495]:.3f} Mpc/h, binned walks {int(dfd.sum())}, "
              f"corrected z range {zc.min():.3f}..{zc.max():.3f} (uncorrected {z.min():.3f}..{z.max():.3f})")


def test_history_correction_errors():
    N = 32
    lp = lib_params(N, nscales=4)
    with D2Z(lp) as d:
        hist = np.linspace(0, 1, 100000).astype(np.float32)
        ch = np.linspace(0, 1, NCHISTORY).astype(np.float32)
        with pytest.raises(D2ZError):  # no result yet
            d.bubble_mfp(hist)
        with pytest.raises(D2ZError):
            d.correct_zreion(hist, ch)
        d.set_delta(field(N))
        d.run()
        with pytest.raises(D2ZError):
            d.bubble_mfp(hist, nskewers=-1)
        # zcur = 49.9 for every fraction: no cell is above it, every skewer
        # stops after its bounded number of draws and the call fails
        with pytest.raises(D2ZError, match="no start cell"):
            d.bubble_mfp(np.ones(100000, np.float32), nskewers=64)
        with pytest.raises(D2ZError):
            d.corrected_history(np.ones(1, np.float32), 0.0, 50.0, np.ones(NX_MFP, np.float32))
        with pytest.raises(D2ZError):
            d.corrected_history(np.ones(8, np.float32), 5.0, 5.0, np.ones(NX_MFP, np.float32))
        # the context still works after the failures
        mfp, _, _ = d.bubble_mfp(D2Z.cumulative_history(d.history(), N), nskewers=256)
        assert np.all(np.isfinite(mfp)) and np.all(mfp > 0)


def test_history_correction_1024():
    """The chain at 1024^3 (the mixed headline run): the same draws and walk
    lengths as the oracle on the full field, the corrected field bit-exact."""
    N, nsk = 1024, 2000
    lp = D2ZParams(N=N, BoxSize=2000.0, Nscales=96, Rmax=120.0, zeta=40.0, M_min=1.4706e8, mode=FUSED_MIXED)
    announce(f"history correction N={N} nskewers={nsk}")
    with D2Z(lp) as d:
        d.generate_delta(1)
        d.run()
        z = d.zreion()
        hist = D2Z.cumulative_history(d.history(), N)
        mfp, mfp_x, dfd = d.bubble_mfp(hist, nskewers=nsk, seed=3)
        mfp_o, x_o, dfd_o, _, _ = ob.bubble_mfp(z, hist, box_mpc(lp), nskewers=nsk, seed=3)
        assert np.array_equal(mfp_x, x_o)
        assert np.abs(dfd.astype(np.int64) - dfd_o.astype(np.int64)).sum() <= 4
        np.testing.assert_allclose(mfp, mfp_o, rtol=1e-6)
        t, fmin, fmax = _fcoll_table()
        ch = d.corrected_history(t, fmin, fmax, mfp)
        assert np.array_equal(ch, ob.corrected_history(t, fmin, fmax, mfp, lp.Rmax, lp.zeta))
        d.correct_zreion(hist, ch)
        zc = d.zreion()
    zc_o = ob.correct_zreion(z, hist, ch)
    assert np.array_equal(zc, zc_o)
    print(f"N={N}: mfp(x=0.5) {mfp[495]:.3f} Mpc/h, binned walks {int(dfd.sum())}")
