"""Multi-box groups: boxes of several sizes run together, one launch per pass
for the boxes of one size on each of four chains (prc_d2z_group_*), through
the C ABI.

* The uneven-box workload: 16 x 128^3 (200 Mpc/h), 16 x 192^3 (300 Mpc/h) and
  16 x 256^3 (400 Mpc/h), distinct seeds, 48 radii each (one cell to 50
  Mpc/h).  Every box's zreion from the group run equals its own prc_d2z_run
  bit for bit (the group runs the same device code per box); one box of each
  size against the FP64 oracle at the FP32 tolerance (max |dz| <= 2e-4, no
  cell above 0.01, x_HII within 1e-3).
* Bubble radii recorded, and errors: unsupported sizes, mismatched radius
  counts.
"""
import numpy as np
import pytest

import d2z_fp64 as oracle
from _cases import oracle_params, zreion_metrics
from conftest import announce
from prc_d2z import D2Z, D2ZError, D2ZParams, FUSED_FP32, Group

pytestmark = pytest.mark.gpu

SIZES = [(128, 200.0), (192, 300.0), (256, 400.0)]


def _params(N, box, nscales=48, rbub=0):
    return D2ZParams(N=N, BoxSize=box, Nscales=nscales, Rmax=50.0, zeta=40.0, M_min=1.4706e8, mode=FUSED_FP32,
                     RecordBubbleSizes=rbub)


def _oracle_zreion(delta, lp, workers=16):
    """oracle.delta2zreion with threaded FFTs (the 256^3 box has 48 radii)."""
    op = oracle_params(lp)
    sc = oracle.setup_scalars(op)
    dhat = oracle.forward(delta, sc.DInit, workers=workers)
    kmag = oracle.kmag_grid(lp.N, sc.box_mpc)
    zre = np.zeros(delta.shape, np.float64)
    for r in range(lp.Nscales):
        z = oracle.barrier(oracle.smooth(dhat, kmag, sc.R[r], workers=workers), sc.A, sc.B[r])
        z = z.astype(np.float32).astype(np.float64)
        np.copyto(zre, z, where=z > zre)
    return zre.astype(np.float32)


def test_group_of_48_uneven_boxes():
    boxes, single, seeds = [], [], []
    seed = 1
    announce("48 boxes: single runs")
    for N, box in SIZES:
        for _ in range(16):
            d = D2Z(_params(N, box))
            d.generate_delta(seed)
            d.run()
            single.append(d.zreion())
            boxes.append(d)
            seeds.append(seed)
            seed += 1
    for d in boxes:  # clobber every result before the group run
        d.generate_delta(999)
        d.run()
    for d, sd in zip(boxes, seeds):
        d.generate_delta(sd)
    announce("48 boxes: group run")
    with Group(boxes) as g:
        g.run()
        for k, (d, z) in enumerate(zip(boxes, single)):
            assert np.array_equal(d.zreion(), z), k
        g.run()  # the captured graph replays
        assert all(np.array_equal(d.zreion(), z) for d, z in zip(boxes, single))
    # one box of each size against the FP64 oracle
    for i in (0, 16, 32):
        d = boxes[i]
        lp = d.params
        announce(f"group box {i} N={lp.N} vs oracle")
        zr = _oracle_zreion(d.delta(), lp)
        m = zreion_metrics(single[i], zr)
        print(i, lp.N, m)
        assert m["max_abs"] <= 2e-4 and m["frac_gt_0.01"] == 0.0 and m["xhii_max"] <= 1e-3
    for d in boxes:
        d.close()


def test_group_with_bubble_radii_and_errors():
    boxes = []
    for k, (N, box) in enumerate([(64, 100.0), (128, 200.0), (192, 300.0), (64, 100.0)]):
        d = D2Z(_params(N, box, nscales=12, rbub=1))
        d.generate_delta(10 + k)
        boxes.append(d)
    ref = []
    for d in boxes:
        d.run()
        ref.append((d.zreion(), d.rbubble(), d.radius_table()))
    announce("group with rbubble")
    with Group(boxes) as g:
        g.run()
        for d, (z, rb, tab) in zip(boxes, ref):
            assert np.array_equal(d.zreion(), z)
            assert np.array_equal(d.rbubble(), rb)
            np.testing.assert_allclose(d.radius_table()[:, 2], tab[:, 2], rtol=1e-6)
    with D2Z(_params(512, 1000.0, nscales=12)) as big:
        with pytest.raises(D2ZError, match="N = 512"):
            Group([boxes[0], big])
    with D2Z(_params(64, 100.0, nscales=13)) as other:
        with pytest.raises(D2ZError, match="Nscales"):
            Group([boxes[0], other])
    for d in boxes:
        d.close()


def test_group_of_one_size_and_of_one_box():
    """Seven 64^3 boxes (the chains split one size: 2 + 2 + 2 + 1 boxes) and a
    group of a single box: every zreion equals the box's own run bit for bit,
    and sigma_grid agrees with it to its summation order."""
    boxes = []
    for k in range(7):
        d = D2Z(_params(64, 100.0, nscales=20))
        d.generate_delta(40 + k)
        boxes.append(d)
    ref = []
    for d in boxes:
        d.run()
        ref.append((d.zreion(), d.radius_table()))
    for d in boxes:  # clobber the results
        d.generate_delta(999)
        d.run()
    for k, d in enumerate(boxes):
        d.generate_delta(40 + k)
    announce("group of 7 boxes of 64^3")
    with Group(boxes) as g:
        g.run()
        for k, (d, (z, tab)) in enumerate(zip(boxes, ref)):
            assert np.array_equal(d.zreion(), z), k
            np.testing.assert_allclose(d.radius_table()[:, 2], tab[:, 2], rtol=1e-6)
    announce("group of one box")
    with Group(boxes[3:4]) as g:
        boxes[3].generate_delta(999)
        boxes[3].run()  # another field's zreion, so that a group run that did nothing shows
        assert not np.array_equal(boxes[3].zreion(), ref[3][0])
        boxes[3].generate_delta(43)
        g.run()
        assert np.array_equal(boxes[3].zreion(), ref[3][0])
    for d in boxes:
        d.close()
