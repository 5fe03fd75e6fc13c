"""Python host binding of libprc_d2z (ctypes over the C ABI in include/prc_d2z.h).

Mirrors the reference d2z program's operator interface for the hot path
(marcelo-alvarez/pr-calc, c++/src/main_d2z.C:15-109):

    params = read_parameter_file("param.d2z")      # io_d2z.C:16-162
    apply_command_line(params, ["-z", "40", ...])   # commandline_d2z.C:33-99
    with D2Z(params) as d:                          # FFTW sizing/alloc/plans
        d.set_delta(read_delta_file(...))           # ReadDeltaFile
        d.run()                                     # forward r2c + delta2zreion()
        z = d.zreion()                              # WriteZreionFile's source

The library is required: importing works without it, but constructing a D2Z
raises if ``libprc_d2z.so`` is missing or no GPU is present.  There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .params import (D2ZParams, apply_command_line, read_parameter_file,  # noqa: F401
                     read_delta_file, write_zreion_file)

LIB_PATH = Path(__file__).resolve().parents[1] / "lib" / "libprc_d2z.so"

PLAIN_FP32, FUSED_FP32, FUSED_MIXED, FUSED_HALF = 0, 1, 2, 3
DENSE, PADDED = 0, 1
NHISTORY = 100000
NCHISTORY = 10000
NX_MFP = 1000
NMFP = 200
MAX_KERNELS = 16

STATUS = {
    0: "OK", -1: "BAD_ARG", -2: "UNSUPPORTED_N", -3: "OOM", -4: "HIP",
    -5: "NAN_INPUT", -6: "STATE", -7: "NO_DEVICE",
}


class CParams(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("Omegam", "Omegab", "Omegal", "h", "ns", "sigma8", "w",
                                         "BoxSize", "zInit", "Rmin")] + \
               [("N", C.c_int), ("Nscales", C.c_int), ("RecordBubbleSizes", C.c_int)] + \
               [(n, C.c_float) for n in ("Rmax", "zeta", "M_min")] + \
               [("mode", C.c_int), ("radii_per_launch", C.c_int), ("device", C.c_int)]


class CScalars(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("DInit", "DHigh", "A", "pnorm", "erfizeta", "sigma_min")]


class CRadiusRow(C.Structure):
    _fields_ = [("R_mpch", C.c_float), ("sigma", C.c_float), ("sigma_grid", C.c_float)]


class CRadiusDiag(C.Structure):
    _fields_ = [("R_mpch", C.c_float), ("sigma", C.c_float), ("sigma_grid", C.c_float),
                ("A", C.c_double), ("B", C.c_double)]


class CTuning(C.Structure):
    _fields_ = [("radii_per_launch", C.c_int), ("z_blocks", C.c_int), ("ms_per_radius", C.c_double),
                ("x_plan", C.c_int), ("z_plan", C.c_int), ("x_radii_per_launch", C.c_int)]


PLAN_VECTOR, PLAN_MATRIX = 0, 1


class CTiming(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("nkernels", C.c_int),
                ("name", (C.c_char * 32) * MAX_KERNELS),
                ("ms", C.c_double * MAX_KERNELS), ("launches", C.c_int * MAX_KERNELS)]


EXPORTS = [
    "prc_d2z_abi_version", "prc_d2z_default_params", "prc_d2z_create", "prc_d2z_destroy",
    "prc_d2z_last_error", "prc_d2z_set_delta", "prc_d2z_set_delta_device", "prc_d2z_run",
    "prc_d2z_run_async", "prc_d2z_sync", "prc_d2z_get_zreion", "prc_d2z_get_rbubble",
    "prc_d2z_get_scalars", "prc_d2z_get_radius_table", "prc_d2z_zreion_device", "prc_d2z_history",
    "prc_d2z_smoothed_field", "prc_d2z_displacement", "prc_d2z_power_spectrum", "prc_d2z_set_timing", "prc_d2z_get_timing", "prc_d2z_device_bytes",
    "prc_d2z_generate_delta", "prc_d2z_get_delta", "prc_d2z_set_graph",
    "prc_d2z_set_launch_config", "prc_d2z_get_launch_config", "prc_d2z_autotune", "prc_d2z_run_many",
    "prc_d2z_bubble_mfp", "prc_d2z_corrected_history", "prc_d2z_correct_zreion",
    "prc_d2z_get_radius_diag", "prc_d2z_set_scalar_overrides", "prc_d2z_history_noalloc",
    "prc_d2z_set_plan", "prc_d2z_get_plan",
    "prc_d2z_group_create", "prc_d2z_group_run", "prc_d2z_group_destroy", "prc_d2z_group_last_error",
]

_lib = None


def load_library(path: str | os.PathLike | None = None):
    """Load libprc_d2z.so (raises OSError when it is missing)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise OSError(f"libprc_d2z.so not built at {p}: run __graft_entry__.build()")
    lib = C.CDLL(str(p))
    vp = C.c_void_p
    fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
    sig = {
        "prc_d2z_abi_version": (C.c_int, []),
        "prc_d2z_default_params": (None, [C.POINTER(CParams)]),
        "prc_d2z_create": (C.c_int, [C.POINTER(CParams), C.POINTER(vp)]),
        "prc_d2z_destroy": (None, [vp]),
        "prc_d2z_last_error": (C.c_char_p, [vp]),
        "prc_d2z_set_delta": (C.c_int, [vp, fp, C.c_int]),
        "prc_d2z_set_delta_device": (C.c_int, [vp, vp, C.c_int]),
        "prc_d2z_run": (C.c_int, [vp]),
        "prc_d2z_run_async": (C.c_int, [vp]),
        "prc_d2z_sync": (C.c_int, [vp]),
        "prc_d2z_get_zreion": (C.c_int, [vp, fp, C.c_int]),
        "prc_d2z_get_rbubble": (C.c_int, [vp, fp, C.c_int]),
        "prc_d2z_get_scalars": (C.c_int, [vp, C.POINTER(CScalars)]),
        "prc_d2z_get_radius_table": (C.c_int, [vp, C.POINTER(CRadiusRow), C.c_int]),
        "prc_d2z_zreion_device": (C.c_int, [vp, C.POINTER(vp)]),
        "prc_d2z_history": (C.c_int, [vp, up]),
        "prc_d2z_smoothed_field": (C.c_int, [vp, C.c_int, fp]),
        "prc_d2z_displacement": (C.c_int, [vp, C.c_int, fp, C.c_int]),
        "prc_d2z_power_spectrum": (C.c_int, [vp, C.c_int, C.c_float, C.c_float, dp, dp, dp,
                                             np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")]),
        "prc_d2z_set_timing": (C.c_int, [vp, C.c_int]),
        "prc_d2z_get_timing": (C.c_int, [vp, C.POINTER(CTiming)]),
        "prc_d2z_device_bytes": (C.c_size_t, [vp]),
        "prc_d2z_generate_delta": (C.c_int, [vp, C.c_uint64, C.c_int]),
        "prc_d2z_get_delta": (C.c_int, [vp, fp, C.c_int]),
        "prc_d2z_set_graph": (C.c_int, [vp, C.c_int]),
        "prc_d2z_set_launch_config": (C.c_int, [vp, C.c_int, C.c_int]),
        "prc_d2z_get_launch_config": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                C.POINTER(C.c_int)]),
        "prc_d2z_autotune": (C.c_int, [vp, C.POINTER(CTuning)]),
        "prc_d2z_run_many": (C.c_int, [C.POINTER(vp), C.c_int]),
        "prc_d2z_bubble_mfp": (C.c_int, [vp, fp, C.c_int, C.c_uint64, fp, fp, fp]),
        "prc_d2z_corrected_history": (C.c_int, [vp, fp, C.c_int, C.c_float, C.c_float, fp, fp]),
        "prc_d2z_correct_zreion": (C.c_int, [vp, fp, fp]),
        "prc_d2z_get_radius_diag": (C.c_int, [vp, C.POINTER(CRadiusDiag), C.c_int]),
        "prc_d2z_set_scalar_overrides": (C.c_int, [vp, vp, vp]),
        "prc_d2z_history_noalloc": (C.c_int, [vp, C.c_int64, up]),
        "prc_d2z_set_plan": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "prc_d2z_get_plan": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "prc_d2z_group_create": (C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(vp)]),
        "prc_d2z_group_run": (C.c_int, [vp]),
        "prc_d2z_group_destroy": (None, [vp]),
        "prc_d2z_group_last_error": (C.c_char_p, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


class D2ZError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"prc_d2z {STATUS.get(code, code)} ({code}): {msg}")
        self.code = code


def _to_cparams(p: D2ZParams) -> CParams:
    c = CParams()
    for name, _ in CParams._fields_:
        setattr(c, name, getattr(p, name))
    return c


class D2Z:
    """One box on one GPU: the reference's delta2zreion operator."""

    def __init__(self, params: D2ZParams):
        self.lib = load_library()
        self.params = params
        self._ctx = C.c_void_p()
        cp = _to_cparams(params)
        rc = self.lib.prc_d2z_create(C.byref(cp), C.byref(self._ctx))
        if rc:
            raise D2ZError(rc, self.lib.prc_d2z_last_error(None).decode())
        self.N = params.N

    # ---------------------------------------------------------------- utils
    def _check(self, rc: int):
        if rc:
            raise D2ZError(rc, self.lib.prc_d2z_last_error(self._ctx).decode())

    def close(self):
        if self._ctx:
            self.lib.prc_d2z_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ interface
    def set_delta(self, delta: np.ndarray):
        N = self.N
        a = np.ascontiguousarray(delta, dtype=np.float32)
        if a.shape == (N, N, N):
            layout = DENSE
        elif a.shape == (N, N, N + 2):
            layout = PADDED
        elif a.size == N ** 3:
            a, layout = a.reshape(N, N, N), DENSE
        else:
            raise ValueError(f"delta must hold {N}^3 cells (dense) or {N}x{N}x{N + 2} (padded)")
        self._check(self.lib.prc_d2z_set_delta(self._ctx, a, layout))

    def set_delta_device(self, ptr: int, layout: int = DENSE):
        self._check(self.lib.prc_d2z_set_delta_device(self._ctx, C.c_void_p(ptr), layout))

    def generate_delta(self, seed: int, white_noise_only: bool = False):
        """Synthetic GRF on the device (Philox normals, BBKS spectrum)."""
        self._check(self.lib.prc_d2z_generate_delta(self._ctx, seed, int(white_noise_only)))

    def delta(self, layout: int = DENSE) -> np.ndarray:
        N = self.N
        out = np.empty((N, N, N + 2) if layout == PADDED else (N, N, N), dtype=np.float32)
        self._check(self.lib.prc_d2z_get_delta(self._ctx, out, layout))
        return out

    def run(self):
        self._check(self.lib.prc_d2z_run(self._ctx))

    def run_async(self):
        self._check(self.lib.prc_d2z_run_async(self._ctx))

    def sync(self):
        self._check(self.lib.prc_d2z_sync(self._ctx))

    def zreion(self, layout: int = DENSE) -> np.ndarray:
        N = self.N
        out = np.empty((N, N, N + 2) if layout == PADDED else (N, N, N), dtype=np.float32)
        self._check(self.lib.prc_d2z_get_zreion(self._ctx, out, layout))
        return out

    def rbubble(self, layout: int = DENSE) -> np.ndarray:
        N = self.N
        out = np.empty((N, N, N + 2) if layout == PADDED else (N, N, N), dtype=np.float32)
        self._check(self.lib.prc_d2z_get_rbubble(self._ctx, out, layout))
        return out

    def scalars(self) -> dict:
        s = CScalars()
        self._check(self.lib.prc_d2z_get_scalars(self._ctx, C.byref(s)))
        return {n: np.float32(getattr(s, n)) for n, _ in CScalars._fields_}

    def radius_table(self) -> np.ndarray:
        n = self.params.Nscales
        rows = (CRadiusRow * n)()
        self._check(self.lib.prc_d2z_get_radius_table(self._ctx, rows, n))
        return np.array([(r.R_mpch, r.sigma, r.sigma_grid) for r in rows], dtype=np.float32)

    def radius_diag(self) -> dict:
        """The per-radius table with the barrier constants: arrays R_mpch, sigma,
        sigma_grid (float32), A, B (float64; B is NaN where sigma(R) > sigma_min)."""
        n = self.params.Nscales
        rows = (CRadiusDiag * n)()
        self._check(self.lib.prc_d2z_get_radius_diag(self._ctx, rows, n))
        out = {f: np.array([getattr(r, f) for r in rows]) for f, _ in CRadiusDiag._fields_}
        for f in ("R_mpch", "sigma", "sigma_grid"):
            out[f] = out[f].astype(np.float32)
        return out

    def set_scalar_overrides(self, sigma=None, erfizeta=None):
        """Pin sigma_min (sigma[0]), sigma(R_r) (sigma[1 + r]) and/or erfizeta to
        given float values (e.g. the reference's GSL results)."""
        sp = None
        if sigma is not None:
            arr = np.ascontiguousarray(sigma, dtype=np.float32)
            if arr.shape != (self.params.Nscales + 1,):
                raise ValueError(f"sigma must hold Nscales + 1 = {self.params.Nscales + 1} values")
            self._sigma_override = arr
            sp = arr.ctypes.data_as(C.c_void_p)
        ep = None
        if erfizeta is not None:
            self._erfizeta_override = np.array([erfizeta], dtype=np.float32)
            ep = self._erfizeta_override.ctypes.data_as(C.c_void_p)
        self._check(self.lib.prc_d2z_set_scalar_overrides(self._ctx, sp, ep))

    def zreion_device_ptr(self) -> int:
        """Device address of the context's dense zreion field."""
        p = C.c_void_p()
        self._check(self.lib.prc_d2z_zreion_device(self._ctx, C.byref(p)))
        return int(p.value)

    def history(self) -> np.ndarray:
        out = np.zeros(NHISTORY, dtype=np.uint32)
        self._check(self.lib.prc_d2z_history(self._ctx, out))
        return out

    @staticmethod
    def cumulative_history(counts: np.ndarray, N: int) -> np.ndarray:
        """GetHistory's float history (bubblehistory_d2z.C:31-32): counts / N^3
        rounded to float, then the running float sum."""
        h = (np.asarray(counts, dtype=np.float64) / float(N) ** 3).astype(np.float32)
        return np.cumsum(h, dtype=np.float32)

    def bubble_mfp(self, history: np.ndarray, nskewers: int = 0, seed: int = 1):
        """GetBubbleMFP (bubblehistory_d2z.C:82-242) on the GPU: mfp[NX_MFP],
        mfp_x[NX_MFP], dfdmfp[NX_MFP, NMFP]."""
        mfp = np.empty(NX_MFP, dtype=np.float32)
        mfp_x = np.empty(NX_MFP, dtype=np.float32)
        dfdmfp = np.empty((NX_MFP, NMFP), dtype=np.float32)
        self._check(self.lib.prc_d2z_bubble_mfp(self._ctx, np.ascontiguousarray(history, dtype=np.float32),
                                                nskewers, seed, mfp, mfp_x, dfdmfp))
        return mfp, mfp_x, dfdmfp

    def corrected_history(self, fcoll: np.ndarray, fmin: float, fmax: float, mfp: np.ndarray) -> np.ndarray:
        """GetCHistory (bubblehistory_d2z.C:44-80): chistory[NCHISTORY]."""
        t = np.ascontiguousarray(fcoll, dtype=np.float32)
        out = np.empty(NCHISTORY, dtype=np.float32)
        self._check(self.lib.prc_d2z_corrected_history(self._ctx, t, len(t), fmin, fmax,
                                                       np.ascontiguousarray(mfp, dtype=np.float32), out))
        return out

    def correct_zreion(self, history: np.ndarray, chistory: np.ndarray):
        """CorrectHistory (bubblehistory_d2z.C:301-343) on the context's zreion."""
        self._check(self.lib.prc_d2z_correct_zreion(self._ctx, np.ascontiguousarray(history, dtype=np.float32),
                                                    np.ascontiguousarray(chistory, dtype=np.float32)))

    def smoothed_field(self, r: int) -> np.ndarray:
        N = self.N
        out = np.empty((N, N, N), dtype=np.float32)
        self._check(self.lib.prc_d2z_smoothed_field(self._ctx, r, out))
        return out

    def displacement(self, axis: int, layout: int = DENSE) -> np.ndarray:
        """Linear displacement field component axis (0, 1, 2) of the current
        density field, as Delta2Velocity (delta2velocity_map.C:8-97)."""
        N = self.N
        out = np.empty((N, N, N + 2) if layout == PADDED else (N, N, N), dtype=np.float32)
        self._check(self.lib.prc_d2z_displacement(self._ctx, axis, out, layout))
        return out

    def power_spectrum(self, nk: int, kmin: float, kmax: float) -> dict:
        """Binned power spectrum of the density field, as PowerSpectrum
        (powerspectrum_lmb.C:8-139): arrays ps, ook2, kmean, nmodes."""
        ps, ook2, kmean = (np.empty(nk, dtype=np.float64) for _ in range(3))
        nmodes = np.empty(nk, dtype=np.int64)
        self._check(self.lib.prc_d2z_power_spectrum(self._ctx, nk, kmin, kmax, ps, ook2, kmean, nmodes))
        return {"ps": ps, "ook2": ook2, "kmean": kmean, "nmodes": nmodes}

    def set_launch_config(self, radii_per_launch: int, z_blocks: int):
        self._check(self.lib.prc_d2z_set_launch_config(self._ctx, radii_per_launch, z_blocks))

    def launch_config(self) -> dict:
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._check(self.lib.prc_d2z_get_launch_config(self._ctx, C.byref(a), C.byref(b), C.byref(c)))
        return {"radii_per_launch": a.value, "z_blocks": b.value, "max_radii": c.value}

    def set_plan(self, x_plan: int, z_plan: int, x_radii_per_launch: int = 16):
        """FFT plans of the fused x and z passes (PLAN_VECTOR / PLAN_MATRIX)."""
        self._check(self.lib.prc_d2z_set_plan(self._ctx, x_plan, z_plan, x_radii_per_launch))

    def plan(self) -> dict:
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._check(self.lib.prc_d2z_get_plan(self._ctx, C.byref(a), C.byref(b), C.byref(c)))
        return {"x_plan": a.value, "z_plan": b.value, "x_radii_per_launch": c.value}

    def autotune(self, cache: str | os.PathLike | None = None) -> dict:
        """Pick FFT plans and launch shape with the library's autotuner.  With
        ``cache`` (a JSON file) a stored choice for this (device, N, mode) is
        reused, and a new one is stored."""
        import json
        key = f"{self.params.device}:N{self.N}:mode{self.params.mode}"
        if cache is not None and Path(cache).exists():
            table = json.loads(Path(cache).read_text())
            if key in table:
                t = table[key]
                self.set_plan(t["x_plan"], t["z_plan"], t["x_radii_per_launch"])
                self.set_launch_config(t["radii_per_launch"], t["z_blocks"])
                return dict(t, cached=True)
        t = CTuning()
        self._check(self.lib.prc_d2z_autotune(self._ctx, C.byref(t)))
        res = {"radii_per_launch": t.radii_per_launch, "z_blocks": t.z_blocks, "ms_per_radius": t.ms_per_radius,
               "x_plan": t.x_plan, "z_plan": t.z_plan, "x_radii_per_launch": t.x_radii_per_launch}
        if cache is not None:
            table = json.loads(Path(cache).read_text()) if Path(cache).exists() else {}
            table[key] = res
            Path(cache).write_text(json.dumps(table, indent=1, sort_keys=True))
        return dict(res, cached=False)

    def set_graph(self, on: bool = True):
        self._check(self.lib.prc_d2z_set_graph(self._ctx, int(on)))

    def set_timing(self, on: bool = True):
        self._check(self.lib.prc_d2z_set_timing(self._ctx, int(on)))

    def timing(self) -> dict:
        t = CTiming()
        self._check(self.lib.prc_d2z_get_timing(self._ctx, C.byref(t)))
        ker = {t.name[k].value.decode(): (t.ms[k], t.launches[k]) for k in range(t.nkernels)}
        return {"total_ms": t.total_ms, "kernels": ker}

    def device_bytes(self) -> int:
        return int(self.lib.prc_d2z_device_bytes(self._ctx))


def run_many(boxes) -> None:
    """Run several D2Z contexts at once (each on its own stream)."""
    boxes = list(boxes)
    if not boxes:
        return
    lib = boxes[0].lib
    arr = (C.c_void_p * len(boxes))(*[b._ctx.value for b in boxes])
    rc = lib.prc_d2z_run_many(arr, len(boxes))
    if rc:
        raise D2ZError(rc, "run_many: " + " | ".join(lib.prc_d2z_last_error(b._ctx).decode() for b in boxes))


def history_noalloc(d_zreion: int, ncells: int) -> np.ndarray:
    """GetHistory counts of a zreion field in device memory of the current
    device, through the library's no-allocation entry point."""
    lib = load_library()
    out = np.zeros(NHISTORY, dtype=np.uint32)
    rc = lib.prc_d2z_history_noalloc(C.c_void_p(d_zreion), int(ncells), out)
    if rc:
        raise D2ZError(rc, lib.prc_d2z_last_error(None).decode())
    return out


class Group:
    """Several boxes (N = 64..256, fused FP32 kernels) run together
    (prc_d2z_group_*): four chains balanced by cells, one launch per pass for
    the boxes of one size on a chain, one graph per step.  The boxes stay
    usable as D2Z objects; after run() each holds its result."""

    def __init__(self, boxes):
        self.lib = load_library()
        self.boxes = list(boxes)
        arr = (C.c_void_p * len(self.boxes))(*[b._ctx.value for b in self.boxes])
        self._g = C.c_void_p()
        rc = self.lib.prc_d2z_group_create(arr, len(self.boxes), C.byref(self._g))
        if rc:
            raise D2ZError(rc, self.lib.prc_d2z_group_last_error(None).decode())

    def run(self):
        rc = self.lib.prc_d2z_group_run(self._g)
        if rc:
            raise D2ZError(rc, self.lib.prc_d2z_group_last_error(self._g).decode())

    def close(self):
        if self._g:
            self.lib.prc_d2z_group_destroy(self._g)
            self._g = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
