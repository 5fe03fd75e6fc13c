"""Mutation check of the oracle's published fitting constants.

For each constant of the oracle's Eisenstein-Hu transfer function
(``_EH.transeh``, c++/src/cosmology.C:290-309), its growth factor (``growth``,
Carroll-Press-Turner and Ma et al. fits, cosmology.C:174-212) and its collapse
threshold (``deltac``, cosmology.C:141-172), this script writes a copy of
oracle/d2z_fp64.py with that one constant scaled by 1.01 and runs
tests/test_oracle_pins.py against the copy.  A pin suite that pins the
constants fails on every mutant; a mutant that survives is printed.

    python tests/mutate_oracle_constants.py

Not collected by pytest (no test_ prefix): it runs one pytest process per
mutant, about a minute in all.
"""
from __future__ import annotations

import re
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
ORACLE = ROOT / "oracle" / "d2z_fp64.py"

# (function, literal as written in the oracle, scale)
MUTANTS = [
    ("transeh", "0.328", None), ("transeh", "431", None), ("transeh", "0.38", None), ("transeh", "22.3", None),
    ("transeh", "44.5", None), ("transeh", "9.83", None), ("transeh", "10 *", "10.1 *"),
    ("transeh", "0.75", None), ("transeh", "0.43", None), ("transeh", "** 4", "** 4.04"),
    ("transeh", "2.728", None), ("transeh", "/ 2.7)", "/ 2.727)"), ("transeh", "2 * math.e", "2.02 * math.e"),
    ("transeh", "1.8", None), ("transeh", "14.2", None), ("transeh", "731", None), ("transeh", "62.5", None),
    ("growth", "4.0 / 7.0", "4.04 / 7.0"), ("growth", "f32(70)", "f32(70.7)"),
    ("growth", "0.255", None), ("growth", "0.305", None), ("growth", "0.0027", None),
    ("growth", "0.366", None), ("growth", "0.266", None), ("growth", "0.0700", None),
    ("deltac", "0.0123", None), ("deltac", "f32(12)", "f32(12.12)"), ("deltac", "/ 20", "/ 20.2"),
    ("deltac", "2.0 / 3.0", "2.02 / 3.0"),
]
# 2.5 in growth cancels in D(z) = (g/x)/g0 and cannot change any output: not a mutant


def function_span(src: str, name: str) -> tuple[int, int]:
    m = re.search(rf"^(\s*)def {name}\(", src, re.M)
    if not m:
        raise KeyError(name)
    indent = len(m.group(1))
    start = m.start()
    nxt = re.compile(rf"^\s{{0,{indent}}}(def |class |@)", re.M)
    m2 = nxt.search(src, m.end())
    return start, (m2.start() if m2 else len(src))


def mutate(src: str, func: str, literal: str, repl: str | None) -> str:
    a, b = function_span(src, func)
    body = src[a:b]
    if literal not in body:
        raise ValueError(f"{literal!r} not in {func}")
    if repl is None:
        repl = repr(float(literal) * 1.01)
    i = body.rfind(literal)  # the code, not a docstring mention before it
    return src[:a] + body[:i] + repl + body[i + len(literal):] + src[b:]


RUNNER = """
import importlib.util, sys
spec = importlib.util.spec_from_file_location("d2z_fp64", sys.argv[1])
mod = importlib.util.module_from_spec(spec)
sys.modules["d2z_fp64"] = mod
spec.loader.exec_module(mod)
import pytest
sys.exit(pytest.main(["-x", "-q", "-p", "no:cacheprovider", sys.argv[2]]))
"""


def main() -> int:
    src = ORACLE.read_text()
    survivors = []
    with tempfile.TemporaryDirectory() as td:
        path = Path(td) / "d2z_fp64.py"
        for func, lit, repl in MUTANTS:
            path.write_text(mutate(src, func, lit, repl))
            r = subprocess.run([sys.executable, "-c", RUNNER, str(path), str(ROOT / "tests" / "test_oracle_pins.py")],
                               cwd=ROOT, capture_output=True, text=True)
            failed = re.search(r"FAILED (\S+)", r.stdout)
            status = f"killed by {failed.group(1).split('::')[-1]}" if r.returncode else "SURVIVED"
            print(f"{func:8s} {lit:>12s} x1.01: {status}", flush=True)
            if not r.returncode:
                survivors.append((func, lit))
    print(f"{len(MUTANTS) - len(survivors)} of {len(MUTANTS)} mutants killed")
    return 1 if survivors else 0


if __name__ == "__main__":
    sys.exit(main())
