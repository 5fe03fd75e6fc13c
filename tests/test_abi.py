"""CPU checks of the C ABI: the library loads, exports every symbol the public
header declares, and rejects bad arguments before touching a device."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import prc_d2z
from prc_d2z import CParams, D2Z, D2ZError, D2ZParams

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "prc_d2z.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(prc_d2z_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "prc_d2z_create" in syms and "prc_d2z_run" in syms
    assert len(syms) >= 20


def test_library_exports_every_declared_symbol():
    lib = prc_d2z.load_library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the binding's own list matches the header
    assert sorted(prc_d2z.EXPORTS) == declared_symbols()


def test_params_struct_layout_matches_header():
    # 10 floats, 3 ints, 3 floats, 3 ints
    assert C.sizeof(CParams) == 4 * 19
    lib = prc_d2z.load_library()
    p = CParams()
    lib.prc_d2z_default_params(C.byref(p))
    assert (p.Rmax, p.zeta, p.M_min) == (256.0, 10.0, pytest.approx(1e8))   # commandline_d2z.C:47-49
    assert p.N == 128 and p.Nscales == 32 and p.mode == 1


def test_abi_version():
    assert prc_d2z.load_library().prc_d2z_abi_version() == 1


@pytest.mark.parametrize("N", [0, 100, 24, 4096])
def test_unsupported_n_rejected_before_device(N):
    with pytest.raises(D2ZError) as ei:
        D2Z(D2ZParams(N=N))
    assert ei.value.code == -2


def test_bad_arguments_rejected():
    for kw in ({"Nscales": 1}, {"h": 0.0}, {"zeta": 1.0}, {"radii_per_launch": 33}, {"mode": 7}):
        with pytest.raises(D2ZError) as ei:
            D2Z(D2ZParams(N=32, **kw))
        assert ei.value.code == -1, kw


def test_null_context_calls_fail_cleanly():
    lib = prc_d2z.load_library()
    assert lib.prc_d2z_run(None) == -1
    assert lib.prc_d2z_sync(None) == -1
    lib.prc_d2z_destroy(None)
    assert lib.prc_d2z_device_bytes(None) == 0


def test_history_correction_entry_points_reject_null():
    lib = prc_d2z.load_library()
    h = np.zeros(prc_d2z.NHISTORY, np.float32)
    mfp = np.ones(prc_d2z.NX_MFP, np.float32)
    ch = np.zeros(prc_d2z.NCHISTORY, np.float32)
    dfd = np.zeros(prc_d2z.NX_MFP * prc_d2z.NMFP, np.float32)
    assert lib.prc_d2z_bubble_mfp(None, h, 0, 1, mfp, mfp, dfd) == -1
    assert lib.prc_d2z_corrected_history(None, mfp, len(mfp), 0.0, 50.0, mfp, ch) == -1
    assert lib.prc_d2z_correct_zreion(None, h, ch) == -1
