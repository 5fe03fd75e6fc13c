"""GPU parity of the linear displacement field (Delta2Velocity,
c++/src/delta2velocity_map.C:8-97) through the C ABI, against the FP64 oracle.

Tolerance: FP32 transforms against FP64, max |d psi| <= 2e-5 * rms(psi) per
component; the plane-wave case against its closed form to the same bound.
"""
import numpy as np
import pytest

import d2z_fp64 as oracle
from _cases import field, lib_params
from conftest import announce
from prc_d2z import D2Z, D2ZError, FUSED_MIXED, FUSED_FP32, PADDED, PLAIN_FP32

pytestmark = pytest.mark.gpu

TOL = 2e-5


def _rel_err(got, ref):
    return float(np.max(np.abs(got.astype(np.float64) - ref)) / np.sqrt(np.mean(ref ** 2)))


@pytest.mark.parametrize("N,mode", [(32, FUSED_FP32), (64, PLAIN_FP32), (128, FUSED_MIXED), (192, FUSED_FP32)])
def test_displacement_vs_oracle(N, mode):
    lp = lib_params(N, mode=mode)
    delta = field(N)
    announce(f"displacement N={N} mode={mode}")
    with D2Z(lp) as d:
        d.set_delta(delta)
        for axis in range(3):
            psi = d.displacement(axis)
            ref = oracle.displacement(delta, lp.BoxSize, axis)
            err = _rel_err(psi, ref)
            print(f"N={N} axis={axis} rel max err {err:.2e}")
            assert err <= TOL, (axis, err)


def test_displacement_plane_wave_closed_form():
    """delta = cos(k.x) with the x Nyquist in the mix: psi_x is zero (own
    Nyquist plane), psi_y and psi_z are -(k_a / k^2) sin(k.x)."""
    N = 64
    lp = lib_params(N)
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    n = (N // 2, 3, -5)
    ph = 2.0 * np.pi * (n[0] * i + n[1] * j + n[2] * k) / N
    delta = np.cos(ph).astype(np.float32)
    dk = 2.0 * np.pi / lp.BoxSize
    kv = np.array(n, dtype=np.float64) * dk
    announce("displacement plane wave N=64")
    with D2Z(lp) as d:
        d.set_delta(delta)
        psi = [d.displacement(a) for a in range(3)]
    assert np.max(np.abs(psi[0])) <= 1e-6 * np.max(np.abs(kv[1] / np.dot(kv, kv)))
    for a in (1, 2):
        ref = -(kv[a] / np.dot(kv, kv)) * np.sin(ph)
        assert _rel_err(psi[a], ref) <= TOL, a


def test_displacement_layout_state_and_errors():
    """Padded output equals dense; the d2z result is unchanged by a
    displacement call in between; a bad axis is rejected."""
    N = 32
    lp = lib_params(N)
    announce("displacement layout/state N=32")
    with D2Z(lp) as d:
        d.set_delta(field(N))
        d.run()
        z0 = d.zreion()
        dense = d.displacement(2)
        padded = d.displacement(2, layout=PADDED)
        assert np.array_equal(padded[:, :, :N], dense)
        d.run()
        assert np.array_equal(d.zreion(), z0)
        with pytest.raises(D2ZError) as ei:
            d.displacement(3)
        assert ei.value.code == -1
