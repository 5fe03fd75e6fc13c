"""ctypes wrapper of the C restatement of the zreion history correction
(bubble_cpu.c: GetBubbleMFP, GetCHistory, zreioncorrect and CorrectHistory of
c++/src/bubblehistory_d2z.C:44-343, Float2Float of c++/src/tables.C:27-38).

TEST INFRASTRUCTURE ONLY: the checker for the library's history-correction
entry points.  The product never imports this file.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import d2z_cpu

NHISTORY = 100000
NCHISTORY = 10000
NX_MFP = 1000
NMFP = 200
NSKEWERS = 20000

_done = False


def _lib():
    global _done
    lib = d2z_cpu.load()
    if not _done:
        fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
        lib.bub_zcur.argtypes = [fp, fp, fp]
        lib.bub_skewers.restype = C.c_int
        lib.bub_skewers.argtypes = [fp, C.c_int, C.c_float, C.c_int, C.c_uint64, fp, dp, up]
        lib.bub_table_value.restype = C.c_float
        lib.bub_table_value.argtypes = [fp, C.c_int, C.c_float, C.c_float, C.c_float]
        lib.bub_mfp_at.restype = C.c_float
        lib.bub_mfp_at.argtypes = [fp, C.c_float]
        lib.bub_chistory.argtypes = [fp, C.c_int, C.c_float, C.c_float, fp, C.c_float, C.c_float, fp]
        lib.bub_zvz.restype = C.c_int
        lib.bub_zvz.argtypes = [fp, fp, fp, C.c_int]
        lib.bub_remap.argtypes = [fp, C.c_longlong, fp]
        _done = True
    return lib


def f32a(x):
    return np.ascontiguousarray(x, dtype=np.float32)


def zcur(history):
    """zcur and x for every ionized fraction of the skewer loop."""
    z = np.empty(NX_MFP, np.float32)
    x = np.empty(NX_MFP, np.float32)
    _lib().bub_zcur(f32a(history), z, x)
    return z, x


def bubble_mfp(zre, history, box, nskewers=NSKEWERS, seed=1):
    """GetBubbleMFP: (mfp, mfp_x, dfdmfp[NX_MFP, NMFP], zcur, sums).  zre is
    the dense (N, N, N) field.  Raises RuntimeError when a skewer finds no
    start cell."""
    zre = f32a(zre)
    N = zre.shape[0]
    z, x = zcur(history)
    sums = np.zeros(NX_MFP, np.float64)
    cnt = np.zeros(NX_MFP * NMFP, np.uint32)
    rc = _lib().bub_skewers(zre.ravel(), N, float(box), int(nskewers), int(seed), z, sums, cnt)
    if rc:
        raise RuntimeError("no start cell found")
    mfp = (sums / (2.0 * nskewers)).astype(np.float32)
    return mfp, x, cnt.astype(np.float32).reshape(NX_MFP, NMFP), z, sums


def table_value(table, tmin, tmax, v):
    t = f32a(table)
    return float(_lib().bub_table_value(t, len(t), float(tmin), float(tmax), float(v)))


def mfp_at(mfp, x):
    return float(_lib().bub_mfp_at(f32a(mfp), float(x)))


def corrected_history(fcoll, fmin, fmax, mfp, rmfp, zeta):
    """GetCHistory: chistory[NCHISTORY]."""
    t = f32a(fcoll)
    ch = np.empty(NCHISTORY, np.float32)
    _lib().bub_chistory(t, len(t), float(fmin), float(fmax), f32a(mfp), float(rmfp), float(zeta), ch)
    return ch


def zvz_table(history, chistory, filled=True):
    out = np.empty(NHISTORY, np.float32)
    if _lib().bub_zvz(f32a(history), f32a(chistory), out, 1 if filled else 0):
        raise RuntimeError("no bracketing interval")
    return out


def correct_zreion(zre, history, chistory):
    """CorrectHistory on a copy of the dense field."""
    out = np.array(zre, dtype=np.float32, copy=True, order="C")
    _lib().bub_remap(out.ravel(), out.size, zvz_table(history, chistory))
    return out
