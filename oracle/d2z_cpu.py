"""ctypes wrapper of the oracle-side C restatement (libd2z_cpu.so).

TEST INFRASTRUCTURE ONLY: used by tests/ and the cpu_baseline leg of bench.py.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libd2z_cpu.so"
_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)
        lib = C.CDLL(str(LIB))
        fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        lib.d2z_cpu_create.restype = C.c_void_p
        lib.d2z_cpu_create.argtypes = [C.c_int, C.c_float, C.c_int]
        lib.d2z_cpu_destroy.argtypes = [C.c_void_p]
        lib.d2z_cpu_forward.restype = C.c_double
        lib.d2z_cpu_forward.argtypes = [C.c_void_p, fp, C.c_float]
        lib.d2z_cpu_reset.argtypes = [C.c_void_p]
        lib.d2z_cpu_radius.restype = C.c_double
        lib.d2z_cpu_radius.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                       C.POINTER(C.c_double)]
        lib.d2z_cpu_get_zreion.argtypes = [C.c_void_p, fp]
        lib.d2z_cpu_get_smooth.argtypes = [C.c_void_p, fp]
        lib.d2z_cpu_get_spectrum.argtypes = [C.c_void_p, fp]
        lib.d2z_cpu_fft1d.restype = C.c_int
        lib.d2z_cpu_fft1d.argtypes = [C.c_int, C.c_int, fp]
        _lib = lib
    return _lib


class CpuD2Z:
    def __init__(self, N: int, box_mpc: float, nthreads: int = 8):
        self.lib = load()
        self.N = N
        self.h = self.lib.d2z_cpu_create(N, float(box_mpc), nthreads)
        if not self.h:
            raise MemoryError("d2z_cpu_create failed")

    def close(self):
        if self.h:
            self.lib.d2z_cpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def forward(self, delta, DInit) -> float:
        return self.lib.d2z_cpu_forward(self.h, np.ascontiguousarray(delta, dtype=np.float32).ravel(),
                                        float(DInit))

    def radius(self, R, sigma, sigma_min, erfizeta, A):
        sg = C.c_double()
        t = self.lib.d2z_cpu_radius(self.h, float(R), float(sigma), float(sigma_min), float(erfizeta),
                                    float(A), C.byref(sg))
        return t, sg.value

    def zreion(self):
        out = np.empty(self.N ** 3, dtype=np.float32)
        self.lib.d2z_cpu_get_zreion(self.h, out)
        return out.reshape(self.N, self.N, self.N)

    def smooth(self):
        out = np.empty(self.N ** 3, dtype=np.float32)
        self.lib.d2z_cpu_get_smooth(self.h, out)
        return out.reshape(self.N, self.N, self.N)

    def spectrum(self):
        N = self.N
        out = np.empty(N * N * (N + 2), dtype=np.float32)
        self.lib.d2z_cpu_get_spectrum(self.h, out)
        return out.view(np.complex64).reshape(N, N, N // 2 + 1)


def fft1d(x: np.ndarray, sign: int) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex64).copy()
    v = a.view(np.float32)
    if load().d2z_cpu_fft1d(a.size, sign, v):
        raise ValueError("unsupported length")
    return a


def run(delta, p, sc, radii=None, nthreads=8):
    """Full restatement run with the oracle's scalars ``sc``; returns
    (zreion, sigma_grid list)."""
    h = CpuD2Z(p.N, sc.box_mpc, nthreads)
    h.forward(delta, sc.DInit)
    sg = []
    for r in (range(p.Nscales) if radii is None else radii):
        _, s = h.radius(sc.R[r], sc.sigma[r], sc.sigma_min, sc.erfizeta, sc.A)
        sg.append(s)
    z = h.zreion()
    h.close()
    return z, sg


def time_sample(delta, lp, radii, nthreads=16):
    """Time forward + the given radii on the CPU (scalars from the FP64 oracle)."""
    import d2z_fp64 as oracle
    op = oracle.D2ZParams(Omegam=lp.Omegam, Omegab=lp.Omegab, Omegal=lp.Omegal, h=lp.h, ns=lp.ns,
                          sigma8=lp.sigma8, w=lp.w, BoxSize=lp.BoxSize, zInit=lp.zInit, Rmin=lp.Rmin,
                          N=lp.N, Nscales=lp.Nscales, Rmax=lp.Rmax, zeta=lp.zeta, M_min=lp.M_min)
    sc = oracle.setup_scalars(op)
    h = CpuD2Z(lp.N, sc.box_mpc, nthreads)
    t_fwd = h.forward(delta, sc.DInit)
    ts = []
    for r in radii:
        t, _ = h.radius(sc.R[r], sc.sigma[r], sc.sigma_min, sc.erfizeta, sc.A)
        ts.append(t)
    h.close()
    return {"t_forward": t_fwd, "t_radius": sum(ts) / len(ts), "fft": "own radix-2/3/4 Stockham"}
