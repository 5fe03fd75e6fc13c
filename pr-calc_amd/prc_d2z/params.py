"""Parameter file, command line and binary field IO of the d2z program.

Host-side restatement of the reference's configuration surface
(marcelo-alvarez/pr-calc):
* ``read_parameter_file``  c++/src/io_d2z.C:16-162  (tag/value lines, ``%``/``#``
  comments, 14 tags, unknown or repeated tags are errors);
* ``apply_command_line``   c++/src/commandline_d2z.C:33-99 (``-f -r -m -z -l -v -e -h``,
  defaults Rmax 256 Mpc/h, zeta 10, M_min 1e8 Msun);
* ``read_delta_file`` / ``write_zreion_file``  c++/src/io_d2z.C:164-246: dense N^3
  float32, x slowest, no header.
"""
from __future__ import annotations

import getopt
import re
from dataclasses import dataclass, field, fields
from pathlib import Path

import numpy as np

# tag -> (attribute, type) in the order of io_d2z.C:33-87
_TAGS = [
    ("Omegam", float), ("Omegab", float), ("Omegal", float), ("h", float), ("ns", float),
    ("sigma8", float), ("w", float), ("BoxSize", float), ("zInit", float), ("Rmin", float),
    ("N", int), ("Nscales", int), ("DeltaFile", str), ("RecordBubbleSizes", int),
]


@dataclass
class D2ZParams:
    # parameter file (struct Parameter, c++/src/allvars_d2z.h:58-64); BoxSize in Mpc/h
    Omegam: float = 0.31
    Omegab: float = 0.049
    Omegal: float = 0.69
    h: float = 0.68
    ns: float = 0.965
    sigma8: float = 0.81
    w: float = -1.0
    BoxSize: float = 200.0
    zInit: float = 0.0
    Rmin: float = 1.0
    N: int = 128
    Nscales: int = 32
    DeltaFile: str = ""
    RecordBubbleSizes: int = 0
    # command line (struct clParameter, c++/src/allvars_d2z.h:66-70)
    Rmax: float = 256.0
    zeta: float = 10.0
    M_min: float = 1e8
    verbose: int = 0
    everbose: int = 0
    lean: int = 0
    Paramfile: str = ""
    Basename: str = "output/output"
    # execution options of this library
    mode: int = 1
    radii_per_launch: int = 0
    device: int = 0
    errors: list = field(default_factory=list)

    def copy(self, **kw):
        d = {f.name: getattr(self, f.name) for f in fields(self)}
        d["errors"] = list(self.errors)
        d.update(kw)
        return D2ZParams(**d)


class ParameterFileError(ValueError):
    pass


_INT_PREFIX = re.compile(r"\s*[+-]?\d+")
_FLOAT_PREFIX = re.compile(r"\s*[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")


def atoi(s: str) -> int:
    """C atoi: the longest leading integer, 0 if none."""
    m = _INT_PREFIX.match(s)
    return int(m.group(0)) if m else 0


def atof(s: str) -> float:
    """C atof of a decimal number: the longest leading float, 0.0 if none."""
    m = _FLOAT_PREFIX.match(s)
    return float(m.group(0)) if m else 0.0


def read_parameter_file(path: str | Path, params: D2ZParams | None = None, strict: bool = True,
                        write_used_values: bool = False) -> D2ZParams:
    """Tag/value parser of io_d2z.C:16-162.

    A line contributes when it holds at least two whitespace-separated words and
    the first does not start with ``%`` or ``#``.  Each tag may appear once; an
    unknown or repeated tag gives the reference's message
    ``Error in <file>: Tag '<tag>' not allowed or multiple defined.``  With
    ``strict`` those messages (and a missing file) raise; otherwise they are
    collected in ``params.errors`` as the reference prints them and goes on.
    """
    p = params.copy() if params else D2ZParams()
    p.Paramfile = str(path)
    path = Path(path)
    if not path.exists():
        msg = f"Parameter file {path} not found."
        if strict:
            raise FileNotFoundError(msg)
        p.errors.append(msg)
        return p
    remaining = {t: ty for t, ty in _TAGS}
    used = []
    for raw in path.read_text().splitlines():
        words = raw.split()
        if len(words) < 2:
            continue
        tag, val = words[0], words[1]
        if tag[0] in "%#":
            continue
        if tag in remaining:
            ty = remaining.pop(tag)
            if ty is int:
                setattr(p, tag, atoi(val))
                used.append(f"{tag:<35s}{getattr(p, tag)}")
            elif ty is float:
                setattr(p, tag, float(np.float32(atof(val))))
                used.append(f"{tag:<35s}{getattr(p, tag):f}")
            else:
                setattr(p, tag, val)
                used.append(f"{tag:<35s}{val}")
        else:
            p.errors.append(f"Error in {path}: Tag '{tag}' not allowed or multiple defined.")
    if write_used_values:
        Path(str(path) + "-usedvalues").write_text("\n".join(used) + "\n")
    if strict and p.errors:
        raise ParameterFileError("\n".join(p.errors))
    return p


def apply_command_line(p: D2ZParams, argv: list[str]) -> D2ZParams:
    """getopt "hvelf:r:m:z:" of commandline_d2z.C:54-99 (argv excludes the
    program name and the parameter file).  ``-h`` or an unknown option raises
    SystemExit(0) like the reference's usage()."""
    q = p.copy()
    try:
        opts, _ = getopt.getopt(argv, "hvelf:r:m:z:")
    except getopt.GetoptError as e:
        raise SystemExit(0) from e
    for o, a in opts:
        if o == "-h":
            raise SystemExit(0)
        if o == "-f":
            q.Basename = f"output/{a}"
        elif o == "-r":
            q.Rmax = float(np.float32(atof(a)))
        elif o == "-m":
            q.M_min = float(np.float32(atof(a)))
        elif o == "-z":
            q.zeta = float(np.float32(atof(a)))
        elif o == "-v":
            q.verbose = 1
        elif o == "-e":
            q.everbose = 1
        elif o == "-l":
            q.lean = 1
    return q


def read_delta_file(path: str | Path, N: int) -> np.ndarray:
    """Dense N^3 float32, x slowest, no header (io_d2z.C:164-207)."""
    a = np.fromfile(path, dtype=np.float32)
    if a.size != N ** 3:
        raise ValueError(f"{path}: expected {N ** 3} float32 values, found {a.size}")
    return a.reshape(N, N, N)


def write_zreion_file(path: str | Path, zreion: np.ndarray) -> None:
    """Dense N^3 float32, x slowest (io_d2z.C:209-246)."""
    np.ascontiguousarray(zreion, dtype=np.float32).tofile(path)
