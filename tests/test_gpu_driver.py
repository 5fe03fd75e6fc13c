"""End-to-end: the host driver (pr-calc_amd/lib/prc_d2z) on a golden fixture,
with the reference's parameter file, command line and file formats."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import d2z_fp64 as oracle
from conftest import announce

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "pr-calc_amd" / "lib" / "prc_d2z"
FIX = ROOT / "tests" / "golden" / "d2z_n32.npz"

PARAM = """# Cosmological parameters
Omegam        0.31
Omegal        0.69
Omegab        0.049
h             0.68
ns            0.965
sigma8        0.81
w             -1
BoxSize       {box}
N             {N}
zInit         0
DeltaFile     {delta}
Rmin          1.0
Nscales       {ns}
"""


@pytest.mark.parametrize("mode", ["fused", "plain", "mixed"])
def test_driver_end_to_end(tmp_path, mode):
    # FP32 paths to the parity bar; the FP16-storage path to its own (max |dz| 0.02)
    tol_z, tol_tau = (0.02, 2e-3) if mode == "mixed" else (2e-4, 1e-4)
    g = np.load(FIX, allow_pickle=False)
    N = int(g["N"])
    g["delta"].astype(np.float32).tofile(tmp_path / "delta")
    (tmp_path / "param.d2z").write_text(PARAM.format(box=float(g["BoxSize"]), N=N, delta=tmp_path / "delta",
                                                     ns=int(g["Nscales"])))
    announce(f"driver mode={mode}")
    cmd = [str(EXE), "param.d2z", "-f", "run", "-m", str(float(g["M_min"])), "-z", str(float(g["zeta"])),
           "-r", str(float(g["Rmax"])), "-v", "-e", "-M", mode]
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    z = np.fromfile(tmp_path / "output" / "run.zreion", dtype=np.float32).reshape(N, N, N)
    assert np.max(np.abs(z.astype(np.float64) - g["zreion"])) <= tol_z
    rows = re.findall(r"^\s+(\d+)\s+([\d.]+)\s+([\d.]+)\s+([\d.]+)$", r.stdout, re.M)
    assert len(rows) == int(g["Nscales"])
    np.testing.assert_allclose([float(x[2]) for x in rows], g["sigma"], atol=1e-3)
    np.testing.assert_allclose([float(x[3]) for x in rows], g["sigma_grid"], atol=1e-3)
    tau = float(re.search(r"tau = ([\d.]+)", r.stdout).group(1))
    ref_tau = oracle.tau(oracle.history_z(), oracle.history_float(oracle.history_counts(g["zreion"]), N),
                         0.31, 0.049, 0.68)
    assert tau == pytest.approx(float(ref_tau), rel=tol_tau)
    used = (tmp_path / "param.d2z-usedvalues").read_text()
    assert "Nscales" in used and "DeltaFile" in used


def test_driver_rejects_bad_parameter_file(tmp_path):
    (tmp_path / "p.d2z").write_text("N 32\nBogus 1\n")
    r = subprocess.run([str(EXE), "p.d2z", "-v", "-e"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "Tag 'Bogus' not allowed or multiple defined." in r.stdout
