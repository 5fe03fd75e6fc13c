"""bench.py's host side on the CPU: the one-GPU contract, the algorithmic byte
model against the committed rocprofv3 traffic profiles, and the scaling of a
profiled launch's measured bytes to the launch size of a run."""
import json
import subprocess
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, str(ROOT))
import bench  # noqa: E402

N3 = 1024.0 ** 3


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""  # no JSON line


@pytest.mark.parametrize("mode,store", [("mixed", 4), ("fused", 4), ("half", 2)])
def test_traffic_profiles_match_the_byte_model(mode, store):
    """Each committed profile states the algorithmic bytes its counters were
    compared with; they must be the byte model's for the profiled launch
    sizes (16 radii per x launch, 32 per y/z launch), and the measured bytes
    must stay within 5% of them (a kernel re-reading its data would not)."""
    prof = bench.traffic_profile("c3_1024", mode)
    assert prof, f"profiles/traffic_c3_1024_{mode}.json missing"
    sizes = {"fx_filter": 16, "inv_y": 32, "fz_barrier": 32}
    for name, nb in sizes.items():
        e = prof[name]
        assert e["algorithmic_bytes_per_launch"] == pytest.approx(bench.kernel_alg_bytes(name, 1024, nb, store))
        assert 0.95 <= e["ratio"] <= 1.05, (mode, name, e["ratio"])
        assert e["bytes_per_launch"] == pytest.approx(e["ratio"] * e["algorithmic_bytes_per_launch"])
    if mode == "mixed":  # the headline profile also holds the forward passes (whole-step traffic)
        for name in ("fwd_z_r2c", "fwd_y", "fwd_x"):
            assert prof[name]["algorithmic_bytes_per_launch"] == pytest.approx(8 * N3)


def test_measured_bytes_follow_the_launch_size():
    """A profile taken at 16 radii per x launch, used for a run at 8: the
    measured bytes scale with the run's algorithmic bytes, not the profile's."""
    prof = bench.traffic_profile("c3_1024", "mixed")
    e = prof["fx_filter"]
    alg8 = bench.kernel_alg_bytes("fx_filter", 1024, 8, 4)
    alg16 = bench.kernel_alg_bytes("fx_filter", 1024, 16, 4)
    assert alg8 == pytest.approx((4 + 4 * 8) * N3)
    b8 = bench.measured_bytes(e, alg8)
    assert b8 == pytest.approx(e["ratio"] * alg8)
    assert bench.measured_bytes(e, alg16) == pytest.approx(e["bytes_per_launch"])
    assert b8 < 0.6 * e["bytes_per_launch"]
    # no byte model: the profiled bytes; no entry: unknown
    assert bench.measured_bytes({"bytes_per_launch": 123.0}, 0.0) == 123.0
    assert bench.measured_bytes(None, alg8) is None
    assert bench.measured_bytes({}, alg8) is None


def test_traffic_profile_of_a_missing_workload_is_empty():
    assert bench.traffic_profile("no_such_workload", "mixed") == {}
    # the committed profiles are plain JSON with the commands that made them
    for mode in ("mixed", "fused", "half"):
        d = json.loads((ROOT / "profiles" / f"traffic_c3_1024_{mode}.json").read_text())
        about = d["_about"]
        assert "rocprofv3 --pmc FETCH_SIZE" in about["fetch_command"]
        assert "rocprofv3 --pmc WRITE_SIZE" in about["write_command"]
