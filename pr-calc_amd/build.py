"""Build step for the prc_d2z library, its host driver and the oracle's C code.

Compiles in-tree (outputs next to their sources, git-ignored):
  pr-calc_amd/lib/libprc_d2z.so     HIP kernels for gfx950 + the C ABI
  pr-calc_amd/lib/prc_d2z           host driver (parameter file, ICs, .zreion)
  oracle/libd2z_cpu.so              oracle-side C restatement (when present)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PRC_ROCM_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions in any generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    r = subprocess.run([str(c) for c in cmd], cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} (exit {r.returncode})")
    # keep the compiler's own warnings visible but drop the known host-pass note
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if msg.strip():
        sys.stderr.write(msg + "\n")


def build_library(jobs: int = 4) -> Path:
    LIB.mkdir(exist_ok=True)
    objdir = LIB / "obj"
    objdir.mkdir(exist_ok=True)
    srcs = [CSRC / "d2z_kernels.hip", CSRC / "pass_kernels.hip", CSRC / "mfma_kernels.hip", CSRC / "mfma_split_kernels.hip",
            CSRC / "grf_kernels.hip",
            CSRC / "bubble_kernels.hip", CSRC / "prc_d2z_api.cpp"]
    objs = []
    procs = []
    for s in srcs:
        o = objdir / (s.stem + ".o")
        objs.append(o)
        cmd = [HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", s, "-o", o]
        if s.suffix == ".cpp":
            cmd[1:1] = ["-x", "hip"]
        procs.append((cmd, subprocess.Popen([str(c) for c in cmd], stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        print("+", " ".join(str(c) for c in cmd), flush=True)
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")
        failed |= p.returncode != 0
    if failed:
        raise RuntimeError("HIP compilation failed")
    so = LIB / "libprc_d2z.so"
    tmp = LIB / "libprc_d2z.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, so)
    return so


def build_driver() -> Path:
    exe = LIB / "prc_d2z"
    src = PKG / "driver" / "main_d2z.cpp"
    if not src.exists():
        return exe
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", src, "-o", exe, f"-L{LIB}", "-lprc_d2z",
          f"-Wl,-rpath,{LIB}", "-Wl,-rpath,$ORIGIN"])
    return exe


def build_oracle() -> Path | None:
    mk = ROOT / "oracle" / "Makefile"
    if not mk.exists():
        return None
    _run(["make", "-B", "-C", ROOT / "oracle", "-j4"])
    return ROOT / "oracle" / "libd2z_cpu.so"


def build_all() -> None:
    build_library()
    build_driver()
    build_oracle()


if __name__ == "__main__":
    build_all()
