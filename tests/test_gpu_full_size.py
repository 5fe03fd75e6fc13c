"""Full-size (1024^3) checks of the transform reuses through closed forms: the
displacement field and the binned power spectrum of a single plane wave
(Delta2Velocity, c++/src/delta2velocity_map.C:8-97; PowerSpectrum,
c++/src/powerspectrum_lmb.C:8-139).  The smaller sizes are compared with the
FP64 oracle in test_gpu_displacement.py and test_gpu_power_spectrum.py.

Tolerances: displacement max error <= 2e-5 of the reference's rms (as at the
small sizes); the plane wave's bin power within 3e-5 relative (FP32
transforms over 2^30 cells), every other bin below 1e-10 of it.
"""
import math

import numpy as np
import pytest

from _cases import lib_params
from conftest import announce
from prc_d2z import D2Z, FUSED_MIXED

pytestmark = pytest.mark.gpu

N = 1024
WAVE = (5, -3, 7)


def _phase_index(n):
    """(n . x) mod N for every cell, int32, without a float phase array."""
    i = np.arange(N, dtype=np.int64)
    p = ((n[0] * i) % N).astype(np.int32)[:, None, None] + ((n[1] * i) % N).astype(np.int32)[None, :, None]
    p = p + ((n[2] * i) % N).astype(np.int32)[None, None, :]
    np.remainder(p, N, out=p)
    return p


def test_plane_wave_displacement_and_power_1024():
    lp = lib_params(N, mode=FUSED_MIXED)
    p = _phase_index(WAVE)
    tab = 2.0 * np.pi * np.arange(N) / N
    delta = np.cos(tab).astype(np.float32)[p]
    dk = 2.0 * np.pi / lp.BoxSize
    kv = np.array(WAVE, dtype=np.float64) * dk
    k2 = float(np.dot(kv, kv))
    nk = 20
    kmin, kmax = 1.07 * dk, 0.93 * dk * N / 2 * math.sqrt(3.0)
    # keep the wave's bin coordinate clear of an edge
    for _ in range(50):
        u = (math.log(math.sqrt(k2)) - math.log(kmin)) / ((math.log(kmax) - math.log(kmin)) / nk)
        if abs(u - round(u)) > 1e-3:
            break
        kmin *= 1.0003
    b0 = int(u)
    announce(f"plane wave N={N}")
    with D2Z(lp) as d:
        d.set_delta(delta)
        del delta
        psi = d.displacement(2)
        got = d.power_spectrum(nk, kmin, kmax)
    sintab = np.sin(tab)
    amp = kv[2] / k2
    rms = abs(amp) / math.sqrt(2.0)
    err = 0.0
    for s in range(0, N, 64):  # slabs, to bound the host memory of the reference
        ref = -amp * sintab[p[s:s + 64]]
        err = max(err, float(np.max(np.abs(psi[s:s + 64].astype(np.float64) - ref))))
    print(f"N={N}: displacement max error / rms {err / rms:.2e}")
    assert err <= 2e-5 * rms
    d3k = np.float64(np.float32(np.float64(np.float32(np.float32(dk) * N)) ** 3))
    expect = (2 * np.pi) ** 3 * (N ** 3 / 4.0) / d3k / got["nmodes"][b0]
    print(f"N={N}: plane-wave bin {b0} power rel error {got['ps'][b0] / expect - 1:.2e}")
    assert got["ps"][b0] == pytest.approx(expect, rel=3e-5)
    others = [b for b in range(nk) if b != b0 and got["nmodes"][b] > 0]
    assert np.max(got["ps"][others]) <= 1e-10 * got["ps"][b0]
