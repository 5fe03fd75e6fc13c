"""The half path (FUSED_HALF: FP32 arithmetic, FP16 storage of the two
inter-pass intermediates, FP16 matrix-core stages at N = 1024 and in the x
pass at N = 2048) against the FP64 oracle.  The compensated mixed path
(FUSED_MIXED) at 1024^3 is checked in test_gpu_parity.py::test_c3_size_1024
to FP32 tolerances.

Tolerances (stated for the half path): max |dzreion| <= 0.02, the fraction of
cells with |dzreion| > 0.01 at most 1e-3, x_HII(z) within 1e-3; the per-radius
delta_R error is reported and bounded by 2e-3 of the field's rms (FP16 has a
unit roundoff of 4.9e-4 and each value is rounded twice).
"""
import numpy as np
import pytest

import d2z_fp64 as oracle
from _cases import field, lib_params, oracle_params, zreion_metrics
from conftest import announce
from prc_d2z import D2Z, D2ZError, D2ZParams, FUSED_FP32, FUSED_HALF

pytestmark = pytest.mark.gpu


def _where_cells_differ(a, b, tol, what):
    """Print where |a - b| > tol: the cell count, the (x, y) z rows and x planes
    they fall in, and the largest differences with their cells, so that a
    failure names the pass (a z row, a y-z plane or scattered cells)."""
    bad = np.abs(a.astype(np.float64) - b) > tol
    n = int(bad.sum())
    if n == 0:
        return
    x, y, zz = np.nonzero(bad)
    rows = np.unique(x * a.shape[1] + y)
    print(f"{what}: {n} cells above {tol}; {rows.size} z rows, x planes {np.unique(x)[:16].tolist()}, "
          f"y {np.unique(y)[:16].tolist()}, z {np.unique(zz)[:16].tolist()}")
    d = np.abs(a.astype(np.float64) - b)[bad]
    for k in np.argsort(d)[::-1][:8]:
        print(f"  cell ({x[k]}, {y[k]}, {zz[k]}): {a[x[k], y[k], zz[k]]} vs {b[x[k], y[k], zz[k]]}")


# 192 runs the radix-3 plans (z c2r of length 96); every size but 1024 runs the
# VALU z pass with its packed FP16 glue
@pytest.mark.parametrize("N", [32, 64, 128, 192, 256])
def test_half_zreion_vs_oracle(N):
    lp = lib_params(N, nscales=10, mode=FUSED_HALF)
    announce(f"half zreion N={N}")
    with D2Z(lp) as d:
        d.set_delta(field(N))
        d.run()
        z = d.zreion()
        tab = d.radius_table()
        d.run()  # a second run of the same graph gives the same bits
        z_again = d.zreion()
    op = oracle_params(lp)
    sc = oracle.setup_scalars(op)
    zr, sg, _ = oracle.delta2zreion(field(N), op, sc)
    m = zreion_metrics(z, zr)
    print(m)
    _where_cells_differ(z, zr, 0.01, "run 1 vs oracle")
    _where_cells_differ(z_again, z, 0.0, "run 2 vs run 1")
    assert np.array_equal(z_again, z)
    assert m["max_abs"] <= 0.02
    assert m["frac_gt_0.01"] <= 1e-3
    assert m["xhii_max"] <= 1e-3
    np.testing.assert_allclose(tab[:, 2], sg, rtol=2e-3)


def test_half_smoothed_field_error_per_radius():
    N = 128
    lp = lib_params(N, nscales=8, mode=FUSED_HALF)
    op = oracle_params(lp)
    sc = oracle.setup_scalars(op)
    dhat = oracle.forward(field(N), sc.DInit)
    kmag = oracle.kmag_grid(N, sc.box_mpc)
    with D2Z(lp) as d:
        d.set_delta(field(N))
        for r in range(lp.Nscales):
            announce(f"half smoothed field N={N} r={r}")
            g = d.smoothed_field(r).astype(np.float64)
            ref = oracle.smooth(dhat, kmag, sc.R[r])
            rms = np.sqrt(np.mean(ref ** 2))
            err_rms = np.sqrt(np.mean((g - ref) ** 2)) / rms
            err_max = np.max(np.abs(g - ref)) / rms
            print(f"r={r} R={sc.R[r]:.3f} Mpc rel rms err {err_rms:.2e} rel max err {err_max:.2e}")
            assert err_rms <= 2e-3
            assert err_max <= 2e-2


@pytest.mark.slow
def test_half_against_fp32_at_512():
    N = 512
    outs = {}
    for mode in (FUSED_FP32, FUSED_HALF):
        lp = D2ZParams(N=N, BoxSize=1000.0, Nscales=64, Rmax=80.0, zeta=40.0, M_min=1.4706e8, mode=mode)
        announce(f"512^3 mode={mode}")
        with D2Z(lp) as d:
            d.generate_delta(1)
            d.run()
            outs[mode] = (d.zreion(), d.history())
    z32, h32 = outs[FUSED_FP32]
    z16, h16 = outs[FUSED_HALF]
    dz = np.abs(z16.astype(np.float64) - z32)
    x32 = np.cumsum(h32.astype(np.int64)) / z32.size
    x16 = np.cumsum(h16.astype(np.int64)) / z16.size
    print(f"512^3 half vs fp32: max|dz| {dz.max():.3e}  frac>0.01 {np.mean(dz > 0.01):.2e} "
          f"x_HII max diff {np.max(np.abs(x32 - x16)):.2e}")
    assert dz.max() <= 0.02
    assert np.max(np.abs(x32 - x16)) <= 1e-3


def test_half_rbubble_1024():
    """1024^3 with bubble radii recorded, first 8 radii: the half path's
    zreion and rbubble against the fused FP32 path.  A cell's recorded radius
    is the first radius reaching the maximum of z(R) over radii; where z(R) is
    flat over several radii (the field's value hardly changes from one
    smoothing scale to the next, 4.4% apart), the FP16 path's error (~6e-3 in
    z) moves that argmax by one or more radius steps.  The test bounds the
    fraction of such cells and shows that the two paths' z values at a
    mismatched cell agree to the path's tolerance: the radius moved, the
    maximum did not."""
    N = 1024
    out = {}
    for mode in (FUSED_FP32, FUSED_HALF):
        lp = D2ZParams(N=N, BoxSize=2000.0, Nscales=8, Rmax=2000.0 / N * (120.0 / (2000.0 / N)) ** (7 / 95),
                       zeta=40.0, M_min=1.4706e8, mode=mode, RecordBubbleSizes=1)
        announce(f"1024^3 rbubble mode={mode}")
        with D2Z(lp) as d:
            d.generate_delta(1)
            d.run()
            out[mode] = (d.zreion(), d.rbubble())
    (z32, rb32), (z16, rb16) = out[FUSED_FP32], out[FUSED_HALF]
    dz = np.abs(z16.astype(np.float64) - z32)
    # neighbouring radii are 4.4% apart: where z(R) is flat over several radii
    # the winning radius moves with errors far below the mixed path's (about
    # 2% of cells at this size)
    radii = np.unique(np.concatenate([rb32[rb32 > 0], rb16[rb16 > 0]]))
    mis = rb16 != rb32
    mism = float(np.mean(mis))
    idx = {float(r): i for i, r in enumerate(radii)}
    steps = np.abs(np.vectorize(idx.get)(rb16[mis]) - np.vectorize(idx.get)(rb32[mis]))
    hist = np.bincount(steps.astype(np.int64), minlength=8)
    print(f"1024^3 rbubble: half vs fp32 max|dz| {dz.max():.2e}, rbubble mismatch fraction {mism:.2e}, "
          f"radius steps of the mismatches (1..7): {hist[1:].tolist()}, max|dz| there {dz[mis].max():.2e}")
    assert radii.size <= 8
    assert dz.max() <= 0.02
    assert np.all((rb16 == 0) == (z16 == 0))
    assert mism <= 0.05
    assert dz[mis].max() <= 0.02  # the maximum agrees where the radius moved


def test_half_matrix_x_2048():
    """At 2048^3 the half path's x pass runs on the matrix cores (2048 = 16 x
    16 x 8, 4 of the 8 columns of a spectrum tile per block).  The smoothed
    field of the smallest radius through that pass and through the vector pass,
    each against the fused FP32 path's (every 8th x plane): relative rms error
    <= 2e-3 and largest error <= 2e-2 of the rms, the half path's bounds."""
    from prc_d2z import PLAN_MATRIX, PLAN_VECTOR
    N = 2048
    r0 = 4000.0 / N
    mk = lambda mode: D2ZParams(N=N, BoxSize=4000.0, Nscales=4, Rmax=r0 * (150.0 / r0) ** (3 / 127), zeta=40.0,
                                M_min=1.4706e8, mode=mode)
    fields = {}
    for key, mode, xplan in (("fp32", FUSED_FP32, None), ("matrix", FUSED_HALF, PLAN_MATRIX),
                             ("vector", FUSED_HALF, PLAN_VECTOR)):
        announce(f"2048^3 smoothed field, {key} x pass")
        with D2Z(mk(mode)) as d:
            d.generate_delta(1)
            if xplan is not None:
                if xplan == PLAN_MATRIX:  # the default plan at this size; no matrix-core z pass here
                    assert d.plan() == {"x_plan": PLAN_MATRIX, "z_plan": PLAN_VECTOR, "x_radii_per_launch": 16}
                    with pytest.raises(D2ZError):
                        d.set_plan(PLAN_MATRIX, PLAN_MATRIX, 16)
                d.set_plan(xplan, PLAN_VECTOR, 16)
                assert d.plan()["x_plan"] == xplan
            f = d.smoothed_field(0)
            fields[key] = f[::8].astype(np.float64)
            del f
    ref = fields.pop("fp32")
    rms = np.sqrt(np.mean(ref ** 2))
    for key, g in fields.items():
        e = g - ref
        err_rms, err_max = np.sqrt(np.mean(e ** 2)) / rms, np.max(np.abs(e)) / rms
        print(f"2048^3 half path, {key} x pass: delta_R rel rms err {err_rms:.2e}, rel max err {err_max:.2e}")
        assert err_rms <= 2e-3 and err_max <= 2e-2


def test_half_repeated_runs_bit_identical():
    """Regression test for a race between the persistent FP16 y pass's tiles
    (a tile's last LDS gather sunk below the barrier that frees the buffer for
    the next tile): 200 fresh 128^3 contexts, each after a 64^3 one and run
    twice, all give the same bits.  Before the fix about 5% of such runs had
    a few waves of one y tile wrong."""
    N = 128
    announce("half 128^3, 200 fresh contexts x 2 runs")
    ref = None
    for t in range(200):
        with D2Z(lib_params(64, nscales=10, mode=FUSED_HALF)) as d:
            d.set_delta(field(64))
            d.run()
        with D2Z(lib_params(N, nscales=10, mode=FUSED_HALF)) as d:
            d.set_delta(field(N))
            d.run()
            z1 = d.zreion()
            d.run()
            z2 = d.zreion()
        if ref is None:
            ref = z1.copy()
        for k, z in ((1, z1), (2, z2)):
            if not np.array_equal(z, ref):
                _where_cells_differ(z, ref.astype(np.float64), 0.0, f"context {t} run {k} vs the first")
            assert np.array_equal(z, ref), (t, k)
