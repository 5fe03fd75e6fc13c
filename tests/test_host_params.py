"""Host-side configuration surface: parameter file, command line, field files
(c++/src/io_d2z.C:16-246, c++/src/commandline_d2z.C:33-99)."""
import numpy as np
import pytest

from prc_d2z.params import (D2ZParams, ParameterFileError, apply_command_line, atof, atoi,
                            read_delta_file, read_parameter_file, write_zreion_file)

FILLED = """# Cosmological parameters

Omegam        0.31
Omegal        0.69
Omegab        0.049
h             0.68
ns            0.965
sigma8        0.81
w             -1

# Box size, dimension, and redshift of input density field

BoxSize\t\t200
# comoving box size in Mpc/h
N\t\t128
#2048 # dimension of the grid
zInit\t\t0
% a percent comment
DeltaFile\t./ICs/delta
Rmin\t\t1.0
Nscales\t\t32
"""


def test_reads_the_template_format(tmp_path):
    f = tmp_path / "param.d2z"
    f.write_text(FILLED)
    p = read_parameter_file(f, write_used_values=True)
    assert (p.N, p.Nscales, p.BoxSize, p.DeltaFile) == (128, 32, 200.0, "./ICs/delta")
    assert p.Omegam == pytest.approx(0.31) and p.w == -1.0 and p.RecordBubbleSizes == 0
    used = (tmp_path / "param.d2z-usedvalues").read_text().splitlines()
    assert used[0].startswith("Omegam") and "0.310000" in used[0]
    assert len(used) == 13


def test_unknown_and_repeated_tags_are_errors(tmp_path):
    f = tmp_path / "p"
    f.write_text(FILLED + "Bogus 3\nN 64\n")
    with pytest.raises(ParameterFileError) as ei:
        read_parameter_file(f)
    msg = str(ei.value)
    assert f"Error in {f}: Tag 'Bogus' not allowed or multiple defined." in msg
    assert "Tag 'N' not allowed or multiple defined." in msg
    p = read_parameter_file(f, strict=False)     # reference behaviour: report and go on
    assert p.N == 128 and len(p.errors) == 2


def test_missing_file(tmp_path):
    with pytest.raises(FileNotFoundError):
        read_parameter_file(tmp_path / "nope")


def test_c_number_parsing():
    assert atoi("128") == 128 and atoi("12.7") == 12 and atoi("x") == 0 and atoi("-3abc") == -3
    assert atof("2e2") == 200.0 and atof(".5x") == 0.5 and atof("abc") == 0.0


def test_command_line_defaults_and_flags():
    p = apply_command_line(D2ZParams(), [])
    assert (p.Rmax, p.zeta, p.M_min, p.Basename) == (256.0, 10.0, 1e8, "output/output")
    q = apply_command_line(p, ["-f", "run1", "-m", "1.4706e8", "-z", "40", "-r", "120", "-l", "-v"])
    assert q.Basename == "output/run1" and q.zeta == 40.0 and q.Rmax == 120.0
    assert q.M_min == pytest.approx(1.4706e8) and q.lean == 1 and q.verbose == 1
    with pytest.raises(SystemExit):
        apply_command_line(p, ["-h"])
    with pytest.raises(SystemExit):
        apply_command_line(p, ["-q"])


def test_field_files_round_trip(tmp_path):
    a = np.random.default_rng(0).standard_normal((8, 8, 8)).astype(np.float32)
    write_zreion_file(tmp_path / "z", a)
    assert (tmp_path / "z").stat().st_size == 4 * 8 ** 3
    np.testing.assert_array_equal(read_delta_file(tmp_path / "z", 8), a)
    with pytest.raises(ValueError):
        read_delta_file(tmp_path / "z", 9)
