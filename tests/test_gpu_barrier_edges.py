"""GPU parity of the barrier's edge cases, through the C ABI, against the oracle.

* Radii below the M_min mass scale, sigma(R) > sigma_min: the reference's
  sqrt(2 (sigma_min^2 - sigma^2)) is NaN there, its z is NaN and its strict
  `zcur > zreion` never updates (delta2zreion_d2z.C:117,138,211), so such a
  radius contributes nothing.  M_min = 3.6e13 Msun puts the first radii of an
  8-radius, 64^3 / 100 Mpc/h run above sigma_min; the cases put those radii
  in a batch of their own, across a batch boundary and inside a batch with
  valid radii, for every path.
* zInit != 0: delta /= D(zInit) before the forward transform (main_d2z.C:57-59).
* The scalar overrides: sigma_min, sigma(R) and erfizeta pinned by the host
  (delta2zreion_d2z.C:17-26,97-100) rebuild B_R exactly and drive the run.

Tolerances as the other parity tests: FP32 paths max |dz| <= 2e-4, no cell
above 0.01; the FP16-storage path max |dz| <= 0.02, fraction above 0.01
<= 1e-3; x_HII(z) within 1e-3.
"""
import numpy as np
import pytest

import d2z_fp64 as oracle
from _cases import field, lib_params, oracle_params, zreion_metrics
from conftest import announce
from prc_d2z import D2Z, D2ZParams, FUSED_FP32, FUSED_HALF, FUSED_MIXED, PLAIN_FP32

pytestmark = pytest.mark.gpu

N = 64
BIG_MMIN = 3.6e13  # Msun: sigma(R) > sigma_min for the first radii
_cache = {}


def _params(mode, rpl=0, rbub=0, M_min=BIG_MMIN, zInit=0.0):
    lp = lib_params(N, nscales=8, mode=mode, rpl=rpl, rbub=rbub)
    lp.M_min = M_min
    lp.zInit = zInit
    return lp


def _oracle(lp, rbub=False, sc=None):
    key = (lp.M_min, lp.zInit, rbub, None if sc is None else float(sc.sigma[0]))
    if key not in _cache:
        op = oracle_params(lp)
        sc = sc if sc is not None else oracle.setup_scalars(op)
        z, sg, rb = oracle.delta2zreion(field(N), op, sc, want_rbubble=rbub)
        _cache[key] = (sc, z, sg, rb)
    return _cache[key]


def _tol(mode):
    return (0.02, 1e-3) if mode == FUSED_HALF else (2e-4, 0.0)


def test_case_has_radii_above_sigma_min():
    sc = oracle.setup_scalars(oracle_params(_params(FUSED_FP32)))
    above = sc.sigma > sc.sigma_min
    assert above[:2].all() and not above[-3:].any(), (sc.sigma, sc.sigma_min)
    assert np.isnan(sc.B[above]).all()


@pytest.mark.parametrize("mode,rpl", [(PLAIN_FP32, 0), (FUSED_FP32, 0), (FUSED_FP32, 2), (FUSED_FP32, 3),
                                      (FUSED_HALF, 0), (FUSED_HALF, 2)])
@pytest.mark.parametrize("rbub", [0, 1])
def test_radii_above_sigma_min_contribute_nothing(mode, rpl, rbub):
    lp = _params(mode, rpl, rbub)
    announce(f"sigma(R) > sigma_min N={N} mode={mode} rpl={rpl} rbub={rbub}")
    with D2Z(lp) as d:
        d.set_delta(field(N))
        d.run()
        z = d.zreion()
        rb = d.rbubble() if rbub else None
        diag = d.radius_diag()
    sc, zr, _, rbr = _oracle(lp, bool(rbub))
    above = sc.sigma > sc.sigma_min
    assert np.array_equal(np.isnan(diag["B"]), above)
    np.testing.assert_allclose(diag["B"][~above], sc.B[~above], rtol=1e-6)
    m = zreion_metrics(z, zr)
    print(m, "ionized fraction", float(np.mean(z > 0)))
    tol, frac = _tol(mode)
    assert m["max_abs"] <= tol and m["frac_gt_0.01"] <= frac and m["xhii_max"] <= 1e-3
    assert np.mean(z > 0) > 0.2  # the valid radii do raise zreion
    if rbub:
        # no cell records a radius above sigma_min: all recorded radii are >= the
        # smallest valid one
        r_valid = float(np.asarray(sc.R)[~above].min())
        assert rb[rb > 0].min() >= r_valid * (1 - 1e-6)
        if mode != FUSED_HALF:
            assert np.mean(rb != rbr) <= 1e-3


@pytest.mark.parametrize("mode", [PLAIN_FP32, FUSED_FP32, FUSED_HALF])
def test_zinit_rescales_the_input(mode):
    """zInit = 6: the field is divided by D(6) before the transform."""
    lp = _params(mode, M_min=1.4706e8, zInit=6.0)
    announce(f"zInit=6 N={N} mode={mode}")
    with D2Z(lp) as d:
        s = d.scalars()
        d.set_delta(field(N) * np.float32(0.2))  # a field at z = 6 is smaller
        d.run()
        z = d.zreion()
    op = oracle_params(lp)
    sc = oracle.setup_scalars(op)
    assert s["DInit"] == sc.DInit and 0.1 < float(sc.DInit) < 0.2
    zr, _, _ = oracle.delta2zreion(field(N) * np.float32(0.2), op, sc)
    m = zreion_metrics(z, zr)
    print(m)
    tol, frac = _tol(mode)
    assert m["max_abs"] <= tol and m["frac_gt_0.01"] <= frac and m["xhii_max"] <= 1e-3


def test_scalar_overrides_rebuild_the_barrier():
    lp = _params(FUSED_FP32, M_min=1.4706e8)
    op = oracle_params(lp)
    sc = oracle.setup_scalars(op)
    # a sigma table and erfizeta a host would take from the reference's GSL run:
    # here the oracle's values, perturbed, so that they differ from the library's
    sig = np.concatenate([[sc.sigma_min], sc.sigma]).astype(np.float32) * np.float32(1.01)
    erfz = np.float32(sc.erfizeta) * np.float32(0.999)
    announce(f"scalar overrides N={N}")
    with D2Z(lp) as d:
        d.set_delta(field(N))
        d.run()
        z0 = d.zreion()
        d.set_scalar_overrides(sigma=sig, erfizeta=erfz)
        diag = d.radius_diag()
        s = d.scalars()
        d.run()
        z1 = d.zreion()
    assert s["sigma_min"] == sig[0] and s["erfizeta"] == erfz
    assert np.array_equal(diag["sigma"], sig[1:])
    B = np.array([oracle.barrier_B(sig[0], x, erfz) for x in sig[1:]])
    assert np.array_equal(diag["B"], B)  # bit for bit: same float and double steps
    sc2 = oracle.Scalars(DInit=sc.DInit, DHigh=sc.DHigh, A=sc.A, pnorm=sc.pnorm, erfizeta=erfz,
                         sigma_min=sig[0], box_mpc=sc.box_mpc, R=sc.R, sigma=sig[1:], B=B)
    zr, _, _ = oracle.delta2zreion(field(N), op, sc2)
    m = zreion_metrics(z1, zr)
    print(m)
    assert m["max_abs"] <= 2e-4 and m["frac_gt_0.01"] == 0.0
    assert np.max(np.abs(z1 - z0)) > 0.05  # the overrides changed the run


@pytest.mark.parametrize("rbub", [0, 1])
def test_radii_above_sigma_min_split_path_1024(rbub):
    """The split matrix-core z pass (mixed mode, N = 1024 only) with the first
    radii above sigma_min, in one batch with valid radii and across a batch
    boundary (4 radii per launch): against the fused FP32 path, whose handling
    of such radii matches the oracle above."""
    out = {}
    for mode, rpl in ((FUSED_FP32, 0), (FUSED_MIXED, 0), (FUSED_MIXED, 4)):
        lp = D2ZParams(N=1024, BoxSize=1600.0, Nscales=8, Rmax=100.0, zeta=40.0, M_min=1.5e14, mode=mode,
                       radii_per_launch=rpl, RecordBubbleSizes=rbub)
        announce(f"1024^3 sigma(R) > sigma_min mode={mode} rpl={rpl} rbub={rbub}")
        with D2Z(lp) as d:
            d.generate_delta(3)
            diag = d.radius_diag()
            d.run()
            out[(mode, rpl)] = (d.zreion(), d.rbubble() if rbub else None)
    above = np.isnan(diag["B"])
    print("radii above sigma_min:", np.nonzero(above)[0].tolist())
    assert above[:3].all() and not above[-2:].any()
    z32 = out[(FUSED_FP32, 0)][0]
    assert np.mean(z32 > 0) > 0.1
    for key in ((FUSED_MIXED, 0), (FUSED_MIXED, 4)):
        zm, rbm = out[key]
        dz = float(np.max(np.abs(zm.astype(np.float64) - z32)))
        print(key, "max|dz| vs fp32", dz)
        assert dz <= 2e-3
        if rbub:
            r_valid = float(diag["R_mpch"][~above].min()) / 0.68
            assert rbm[rbm > 0].min() >= r_valid * (1 - 1e-5)
