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
    # WriteHistory's four columns (io_d2z.C:284-314); no fcoll table here
    hist = np.loadtxt(tmp_path / "output" / "run.history")
    assert hist.shape == (oracle.NHISTORY, 4)
    assert np.all(np.isnan(hist[:, 2]))
    smin = float(g["scalars"][5])  # sigma_min of the oracle's set-up
    zz = hist[1::997, 0]
    np.testing.assert_allclose(hist[1::997, 3], [oracle.fcoll(z, 0.31, 0.69, -1.0, smin) for z in zz],
                               rtol=1e-4, atol=1e-30)


def test_driver_rejects_bad_parameter_file(tmp_path):
    (tmp_path / "p.d2z").write_text("N 32\nBogus 1\n")
    r = subprocess.run([str(EXE), "p.d2z", "-v", "-e"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "Tag 'Bogus' not allowed or multiple defined." in r.stdout


def test_driver_history_correction(tmp_path):
    """main_d2z.C:78-90: .mfp and .mfppdf written, and the .zreion corrected
    through fcoll_table.tab; without the table the uncorrected field."""
    import bubble as ob

    g = np.load(FIX, allow_pickle=False)
    N, box = int(g["N"]), float(g["BoxSize"])
    g["delta"].astype(np.float32).tofile(tmp_path / "delta")
    (tmp_path / "param.d2z").write_text(PARAM.format(box=box, N=N, delta=tmp_path / "delta",
                                                     ns=int(g["Nscales"])))
    n, tmin, tmax = 1000, 0.0, 50.0
    zt = tmin + (tmax - tmin) / n * np.arange(n)
    table = (0.03 * np.exp(-((zt / 9.0) ** 2))).astype(np.float32)
    with open(tmp_path / "fcoll_table.tab", "wb") as f:
        np.array([tmin, tmax], np.float32).tofile(f)
        np.array([n], np.int32).tofile(f)
        table.tofile(f)
    announce("driver history correction")
    base = [str(EXE), "param.d2z", "-m", str(float(g["M_min"])), "-z", str(float(g["zeta"])),
            "-r", str(float(g["Rmax"])), "-v", "-e", "-M", "fused"]
    r0 = subprocess.run(base + ["-f", "raw", "-T", "missing.tab"], cwd=tmp_path, capture_output=True, text=True,
                        timeout=120)
    assert r0.returncode == 0 and "missing.tab not found" in r0.stdout
    r1 = subprocess.run(base + ["-f", "cor"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r1.stdout[-1500:], r1.stderr[-1500:])
    assert r1.returncode == 0
    raw = np.fromfile(tmp_path / "output" / "raw.zreion", dtype=np.float32).reshape(N, N, N)
    cor = np.fromfile(tmp_path / "output" / "cor.zreion", dtype=np.float32).reshape(N, N, N)
    hist = np.cumsum((oracle.history_counts(raw) / float(N) ** 3).astype(np.float32), dtype=np.float32)
    # the walks measure Mpc: BoxSize/h (main_d2z.C:29)
    box_mpc = float(np.float32(box) / np.float32(0.68))
    mfp, mfp_x, dfd, _, _ = ob.bubble_mfp(raw, hist, box_mpc, nskewers=ob.NSKEWERS, seed=1)
    txt = np.loadtxt(tmp_path / "output" / "cor.mfp")
    assert txt.shape == (ob.NX_MFP, 2)
    np.testing.assert_allclose(txt[:, 0], mfp_x, atol=1e-6)
    np.testing.assert_allclose(txt[:, 1], mfp, rtol=1e-5, atol=2e-6)
    with open(tmp_path / "output" / "cor.mfppdf", "rb") as f:
        h1 = np.fromfile(f, np.int32, 2)
        h2 = np.fromfile(f, np.float32, 4)
        pdf = np.fromfile(f, np.float32)
    assert list(h1) == [ob.NX_MFP, ob.NMFP]
    np.testing.assert_allclose(h2, [0.01, 0.999, np.log10(box_mpc / N), np.log10(box_mpc)], rtol=1e-6)
    assert pdf.size == ob.NX_MFP * ob.NMFP and np.abs(pdf - dfd.ravel()).sum() <= 4
    ch = ob.corrected_history(table, tmin, tmax, mfp, float(g["Rmax"]), float(g["zeta"]))
    ref = ob.correct_zreion(raw, hist, ch)
    assert np.max(np.abs(cor - ref)) < 1e-3
    assert not np.array_equal(cor, raw)
    # main_d2z.C:97-101: GetHistory again on the corrected field, its tau
    # printed, and WriteHistory's four columns from it
    xh = oracle.history_float(oracle.history_counts(cor), N)
    zh = oracle.history_z()
    tau_cor = float(oracle.tau(zh, xh, 0.31, 0.049, 0.68))
    tau_raw = float(oracle.tau(zh, oracle.history_float(oracle.history_counts(raw), N), 0.31, 0.049, 0.68))
    tau = float(re.search(r"tau = ([\d.]+)", r1.stdout).group(1))
    assert abs(tau_cor - tau_raw) > 1e-5  # the correction moves tau
    assert tau == pytest.approx(tau_cor, abs=2e-6)
    hcol = np.loadtxt(tmp_path / "output" / "cor.history")
    assert hcol.shape == (ob.NHISTORY, 4)
    np.testing.assert_allclose(hcol[:, 1], xh, rtol=1e-6, atol=1e-7)
    idx = np.arange(1, ob.NHISTORY, 499)
    np.testing.assert_allclose(hcol[idx, 2], [ob.table_value(table, tmin, tmax, z) for z in zh[idx]], rtol=1e-5,
                               atol=1e-12)
    smin = float(g["scalars"][5])
    np.testing.assert_allclose(hcol[idx, 3], [oracle.fcoll(z, 0.31, 0.69, -1.0, smin) for z in zh[idx]],
                               rtol=1e-4, atol=1e-30)
