"""Shared small-box cases for the parity tests (C1-like cell size 1.5625 Mpc/h)."""
import numpy as np

import d2z_fp64 as oracle
import grf
from prc_d2z import D2ZParams

CELL = 1.5625  # Mpc/h, the 128^3 / 200 Mpc/h cell


def box_for(N: int) -> float:
    return CELL * N


def box_mpc(lp: D2ZParams) -> float:
    """BoxSize in Mpc as the d2z program holds it (main_d2z.C:29, float division)."""
    return float(np.float32(lp.BoxSize) / np.float32(lp.h))


def lib_params(N, nscales=8, mode=1, rpl=0, rbub=0, rmax=None, zeta=40.0):
    box = box_for(N)
    return D2ZParams(N=N, BoxSize=box, Nscales=nscales, Rmax=rmax if rmax else box / 4, zeta=zeta,
                     M_min=1.4706e8, mode=mode, radii_per_launch=rpl, RecordBubbleSizes=rbub)


def oracle_params(lp: D2ZParams) -> oracle.D2ZParams:
    return oracle.D2ZParams(Omegam=lp.Omegam, Omegab=lp.Omegab, Omegal=lp.Omegal, h=lp.h, ns=lp.ns,
                            sigma8=lp.sigma8, w=lp.w, BoxSize=lp.BoxSize, zInit=lp.zInit, Rmin=lp.Rmin,
                            N=lp.N, Nscales=lp.Nscales, RecordBubbleSizes=lp.RecordBubbleSizes,
                            Rmax=lp.Rmax, zeta=lp.zeta, M_min=lp.M_min)


_field_cache = {}


def field(N: int, seed: int = 1) -> np.ndarray:
    key = (N, seed)
    if key not in _field_cache:
        _field_cache[key] = grf.grf(N, box_for(N), seed)
    return _field_cache[key]


def zreion_metrics(z_gpu, z_ref):
    d = np.abs(z_gpu.astype(np.float64) - z_ref.astype(np.float64))
    n = z_ref.size
    xg = oracle.xhii_from_counts(oracle.history_counts(z_gpu), n)
    xr = oracle.xhii_from_counts(oracle.history_counts(z_ref), n)
    return {"max_abs": float(d.max()), "frac_gt_0.01": float(np.mean(d > 0.01)),
            "xhii_max": float(np.max(np.abs(xg - xr)))}
