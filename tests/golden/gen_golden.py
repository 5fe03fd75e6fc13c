"""Writes the golden fixtures of the d2z path from the FP64 oracle.

The reference ships no golden vectors for delta2zreion (its only tests are MPI
stubs), so these fixtures are produced by this repository's oracle on
synthetic fields and frozen: inputs, scalars, the radius table and zreion.

    python tests/golden/gen_golden.py        # rewrites tests/golden/*.npz
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1] / "oracle"))
import d2z_fp64 as o  # noqa: E402
import grf  # noqa: E402

CASES = {
    # name: N, BoxSize [Mpc/h], Nscales, Rmax [Mpc/h], zeta, seed
    "d2z_n16": (16, 25.0, 6, 6.25, 40.0, 1),
    "d2z_n32": (32, 50.0, 8, 12.5, 40.0, 2),
    "d2z_n32_zeta10": (32, 50.0, 5, 12.5, 10.0, 3),
}


def make(name):
    N, box, ns, rmax, zeta, seed = CASES[name]
    p = o.D2ZParams(N=N, BoxSize=box, Nscales=ns, Rmax=rmax, zeta=zeta, M_min=1.4706e8)
    sc = o.setup_scalars(p)
    delta = grf.grf(N, box, seed)
    z, sg, rb = o.delta2zreion(delta, p, sc, want_rbubble=True)
    counts = o.history_counts(z)
    nz = np.nonzero(counts)[0]
    np.savez_compressed(
        HERE / f"{name}.npz", N=N, BoxSize=box, Nscales=ns, Rmax=rmax, zeta=zeta, M_min=1.4706e8,
        seed=seed, delta=delta, zreion=z, rbubble=rb, R=sc.R, sigma=sc.sigma, B=sc.B,
        sigma_grid=np.array(sg), scalars=np.array([sc.DInit, sc.DHigh, sc.A, sc.pnorm, sc.erfizeta,
                                                    sc.sigma_min], dtype=np.float32),
        hist_bins=nz.astype(np.int32), hist_counts=counts[nz].astype(np.int64))


if __name__ == "__main__":
    for n in CASES:
        make(n)
        print("wrote", n)
