#!/usr/bin/env python3
"""Benchmark of the delta2zreion hot path on one MI355X.

One step = one full delta2zreion pass over a box resident in HBM: forward r2c
of the density field, then for every radius the filter, c2r, barrier and
running maximum (the work of c++/src/main_d2z.C:57-71 in
marcelo-alvarez/pr-calc).  Default workload: 1024^3 cells, 2000 Mpc/h box,
96 radii from one cell to 120 Mpc/h, zeta 40, M_min 1e8 Msun/h, input a
synthetic Gaussian random field generated on the GPU (seed 1).  Default path:
the batched mixed-precision fused path; the plain FP32 path is measured beside
it on the same box (plain_fp32, speedup_vs_plain_fp32).

Prints ONE JSON line.  Extra objects:
  roofline      dominant kernel: algorithmic bytes per launch / average launch
                time (HIP events on the library's stream during the timed
                steps) against the 8 TB/s HBM peak
  cpu_baseline  the CPU restatement timed on a bounded sample of the same
                workload on this host
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "pr-calc_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell-radius of the whole path: one half-spectrum
# intermediate written and read per radius, 4 + 4 B in FP32 storage, 2 + 2 B in
# FP16 storage (the half mode)
B_ALG_PATH = {"plain": 8.0, "fused": 8.0, "mixed": 8.0, "half": 4.0}

WORKLOADS = {
    # name: N, BoxSize [Mpc/h], Nscales, Rmax [Mpc/h]
    "c1_128": (128, 200.0, 32, 50.0),
    "c2_512": (512, 1000.0, 64, 80.0),
    "c3_1024": (1024, 2000.0, 96, 120.0),
    "c5_2048": (2048, 4000.0, 128, 150.0),
}
# 48 uneven boxes: 16 x 128^3 (200 Mpc/h), 16 x 192^3 (300), 16 x 256^3 (400),
# distinct seeds, 48 radii each from one cell to 50 Mpc/h
MULTIBOX = [(128, 200.0), (192, 300.0), (256, 400.0)]
MODES = {"plain": 0, "fused": 1, "mixed": 2, "half": 3}
METRIC = "cell\u00b7radius updates/s at 1024\u00b3\u00d796R: batched mixed vs plain FP32; HBM GB/s"
DTYPES = {"plain": "f32", "fused": "f32",
          "mixed": "f32 transforms: radix-16 stages on f16 matrix cores with split hi/lo operands "
                   "(hi*hi + hi*lo + lo*hi) and f32 accumulation; f32 intermediates, f32 vector stages",
          "half": "f16 storage and f16 matrix-core radix-16 stages with f32 accumulation; f32 vector and barrier"}
STORE_BYTES = {"plain": 4, "fused": 4, "mixed": 4, "half": 2}


FX_MAX_RADII = 16  # radii per fused x-pass launch (the library's launch limit)


def kernel_alg_bytes(name: str, N: int, nb: int, store_bytes: int) -> float:
    """Algorithmic HBM bytes one launch of a kernel must move.

    Intermediates hold N^3/2 complex = N^3 * store_bytes/... per radius:
    fx_filter   reads the half spectrum once (4 B/cell) and writes nb
                intermediates (store B/cell each)
    inv_y       reads and writes nb intermediates
    fz_barrier  reads nb intermediates and reads+writes zreion (8 B/cell)
    plain kernels: filter 8, inv_x/inv_y/inv_z_c2r 8, barrier 12, moments 4 B/cell
    """
    cells = float(N) ** 3
    s = float(store_bytes)
    table = {
        "fx_filter": 4.0 + s * nb,
        "inv_y": 2 * s * nb,
        "fz_barrier": s * nb + 8.0,
        "filter": 8.0, "inv_x": 8.0, "inv_z_c2r": 8.0, "barrier": 12.0, "moments": 4.0,
        "fwd_z_r2c": 8.0, "fwd_y": 8.0, "fwd_x": 8.0,
    }
    return table.get(name, 0.0) * cells


def traffic_profile(workload: str, mode: str) -> dict:
    """The committed rocprofv3 counter summary of a workload and mode
    (FETCH_SIZE and WRITE_SIZE per launch, calibrated), or {}."""
    f = ROOT / "profiles" / f"traffic_{workload}_{mode}.json"
    if not f.exists():
        return {}
    try:
        return json.loads(f.read_text())
    except Exception:
        return {}


def measured_bytes(entry, alg_per_launch: float):
    """HBM bytes of one launch at this run's launch size: the profiled
    launch's measured / algorithmic ratio times this launch's algorithmic
    bytes (the profile may hold another number of radii per launch); the
    profiled bytes for a kernel without a byte model; None without an entry."""
    if not entry:
        return None
    r = entry.get("ratio")
    if r is not None and alg_per_launch:
        return r * alg_per_launch
    return entry.get("bytes_per_launch")


def cpu_baseline_sample(delta: np.ndarray, lp, nthreads: int) -> dict:
    """Time the oracle-side CPU restatement on a bounded sample: the forward
    transform plus four radii spread over the schedule (about 14 s at 1024^3
    on 16 threads); the full-run rate is extrapolated as
    N^3 Nscales / (t_forward + Nscales mean(t_radius))."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import d2z_cpu  # oracle-side C restatement (float, OpenMP)

    N, NS = lp.N, lp.Nscales
    radii = sorted({0, NS // 3, (2 * NS) // 3, NS - 1})
    res = d2z_cpu.time_sample(delta, lp, radii=radii, nthreads=nthreads)
    t_fwd, t_rad = res["t_forward"], res["t_radius"]
    rate = N ** 3 * NS / (t_fwd + NS * t_rad)
    return {"value": rate, "unit": "cell*radius/s", "cores": nthreads, "kind": "port",
            "sample": f"{N}^3 forward r2c + {len(radii)} of {NS} radii (r={radii}) with the C FP32 restatement "
                      f"({res['fft']} FFT); t_forward={t_fwd:.2f}s mean t_radius={t_rad:.2f}s, "
                      f"rate extrapolated to {NS} radii",
            "host": host_cpu()}


def host_cpu() -> str:
    """CPU model and logical CPU count of the host the baseline ran on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs visible"


def history_correction_times(d, N):
    """Wall times (ms) of the history correction on the measured field: the
    calls of main_d2z.C:78-90 in order.  fcoll(z) is a synthetic table
    (0.03 exp(-(z/9)^2) on [0, 50], 2000 entries): the real one comes from the
    reference's colossus-based script."""
    times = {}

    def timed(name, f):
        t0 = time.perf_counter()
        r = f()
        times[name] = round((time.perf_counter() - t0) * 1e3, 3)
        return r

    from prc_d2z import D2Z

    counts = timed("history_ms", d.history)
    hist = D2Z.cumulative_history(counts, N)
    mfp, _, dfd = timed("bubble_mfp_ms", lambda: d.bubble_mfp(hist))
    zt = np.arange(2000) * (50.0 / 2000)
    table = (0.03 * np.exp(-((zt / 9.0) ** 2))).astype(np.float32)
    ch = timed("corrected_history_ms", lambda: d.corrected_history(table, 0.0, 50.0, mfp))
    timed("correct_zreion_ms", lambda: d.correct_zreion(hist, ch))
    times["skewer_walks"] = 2 * 20000 * 1000
    times["binned_walks"] = int(dfd.sum())
    return times


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="mixed", choices=list(MODES),
                    help="mixed (default): fused path whose radix-16 stages run on the matrix cores with "
                         "split hi/lo FP16 operands and FP32 accumulation, FP32 intermediates; fused: fused "
                         "path on the vector ALUs in FP32; half: fused path with FP16 intermediates and plain "
                         "FP16 matrix-core stages; plain: one radius per pass in FP32")
    ap.add_argument("--no-plain-compare", action="store_true",
                    help="skip the plain FP32, fused FP32 and half paths measured beside the default one")
    ap.add_argument("--workload", default="c3_1024", choices=list(WORKLOADS) + ["c4_multibox"])
    ap.add_argument("--radii-per-launch", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the library's autotuner (FFT plans of the x and z passes, radii per x launch and "
                         "per batch, z grid), which otherwise runs once before the warm-up")
    ap.add_argument("--plan", default=None, metavar="X,Z",
                    help="fixed FFT plans instead of tuning: vector or matrix for the x and z pass, e.g. matrix,vector")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--history", action="store_true",
                    help="also run the GetHistory histogram once after timing (read-only pass over zreion)")
    ap.add_argument("--history-correction", action="store_true",
                    help="after timing, run the history correction once (GetHistory, bubble mean free "
                         "paths, corrected history from a synthetic fcoll(z) table, remap of zreion) and "
                         "report its wall times under history_correction")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: the delta2zreion path runs the whole grid on a single "
              f"MI355X; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    from prc_d2z import D2Z, D2ZParams

    if args.workload == "c4_multibox":
        multibox(args)
        return
    N, box, NS, rmax = WORKLOADS[args.workload]
    lp = D2ZParams(N=N, BoxSize=box, Nscales=NS, Rmax=rmax, zeta=40.0, M_min=1.4706e8,
                   mode=MODES[args.mode], radii_per_launch=args.radii_per_launch)
    d = D2Z(lp)
    d.generate_delta(seed=1)          # input resident in HBM before timing
    tuning = None
    if args.plan and args.mode != "plain":
        xp, zp = (0 if v.strip() == "vector" else 1 for v in args.plan.split(","))
        d.set_plan(xp, zp, FX_MAX_RADII)
    elif not args.no_tune and args.mode != "plain":
        tuning = d.autotune()        # untimed: plans and launch shape for this box
    for _ in range(args.warmup):
        d.run()
    d.set_timing(True)                # HIP events around every launch on the library's stream
    d.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        d.run_async()
    d.sync()
    t1 = time.perf_counter()
    kt = d.timing()                   # events of the last timed step
    d.set_timing(False)
    elapsed = t1 - t0
    ms_step = elapsed / args.steps * 1e3
    value = N ** 3 * NS * args.steps / elapsed

    kernels = kt["kernels"]
    dom = max(kernels.items(), key=lambda kv: kv[1][0])
    dom_name = dom[0]
    # radii per fused launch as the library runs them (it may lower the
    # request to fit the intermediates in HBM)
    nb = d.launch_config()["radii_per_launch"] if args.mode != "plain" else 1
    plan = d.plan() if args.mode != "plain" else None
    fx_nb = plan["x_radii_per_launch"] if plan else FX_MAX_RADII
    store = STORE_BYTES[args.mode]

    def alg_step_of(name):
        """Algorithmic bytes of a kernel over the step (the last fused batch may
        carry fewer radii; the x pass fills a batch in launches of at most
        x_radii_per_launch, each reading the spectrum)."""
        k_n = kernels[name][1]
        if args.mode == "plain" or name not in ("fx_filter", "inv_y", "fz_barrier"):
            return kernel_alg_bytes(name, N, 1, store) * k_n
        launches = [min(nb, NS - r0) for r0 in range(0, NS, nb)]
        if name == "fx_filter":
            launches = [min(fx_nb, n - o) for n in launches for o in range(0, n, fx_nb)]
        return sum(kernel_alg_bytes(name, N, n, store) for n in launches)

    tprof = traffic_profile(args.workload, args.mode)

    def kernel_roof(name):
        """A kernel's algorithmic bytes over its event time, and its measured
        HBM bytes per launch."""
        k_ms, k_n = kernels[name]
        alg_step = alg_step_of(name)
        achieved = alg_step / (k_ms * 1e-3) / 1e9 if alg_step else None
        return {"achieved": achieved, "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                "traffic": measured_bytes(tprof.get(name), alg_step / k_n),
                "avg_launch_ms": k_ms / k_n, "alg_bytes_per_launch": alg_step / k_n, "ms_per_step": k_ms}

    dr = kernel_roof(dom_name)
    roof = {"bound": "hbm", "kernel": dom_name, "achieved": dr["achieved"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": dr["frac"], "traffic": dr["traffic"],
            "avg_launch_ms": dr["avg_launch_ms"], "alg_bytes_per_launch": dr["alg_bytes_per_launch"],
            "b_alg_path": B_ALG_PATH[args.mode],
            "path_achieved": value * B_ALG_PATH[args.mode] / 1e9,
            "path_frac": value * B_ALG_PATH[args.mode] / 1e9 / HBM_PEAK_GBS}
    # every fused pass, not only the dominant one (in the mixed path the x and
    # y passes take about the same time per step)
    if args.mode != "plain":
        roof["fused_kernels"] = {k: kernel_roof(k) for k in ("fx_filter", "inv_y", "fz_barrier") if k in kernels}
    # measured HBM bytes of the whole step (rocprof FETCH_SIZE + WRITE_SIZE per
    # kernel launch, committed under profiles/) over the step time
    step_bytes = 0.0
    for name, (_, k_n) in kernels.items():
        b = measured_bytes(tprof.get(name), alg_step_of(name) / k_n)
        if b is None:
            step_bytes = None
            break
        step_bytes += b * k_n
    if step_bytes:
        roof["path_traffic_bytes"] = step_bytes
        roof["path_traffic_frac"] = step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS

    # the same steps replayed as the captured hipGraph (no per-launch events),
    # each step timed on its own for the median
    d.run()
    d.sync()
    step_ms = []
    for _ in range(args.steps):
        g0 = time.perf_counter()
        d.run_async()
        d.sync()
        step_ms.append((time.perf_counter() - g0) * 1e3)
    graph_ms = sum(step_ms) / len(step_ms)

    if args.history:
        h = d.history()
        print(f"history: {int(h.sum())} cells counted", file=sys.stderr)
    hc = None
    if args.history_correction:
        hc = history_correction_times(d, N)
    plain, others = None, {}
    if args.mode != "plain" and not args.no_plain_compare:
        plain = plain_fp32_rate(lp, steps=min(args.steps, 2))
        # the other fused paths on the same box, timed the same way
        for other in ("fused", "mixed", "half"):
            if other != args.mode:
                others[other] = path_rate(lp, other, steps=min(args.steps, 3), tune=not args.no_tune)
    cpu = None
    if not args.no_cpu_baseline:
        try:
            delta = d.delta()
            cpu = cpu_baseline_sample(delta, lp, args.cpu_threads)
        except Exception as e:  # the baseline is reported, not required
            cpu = {"value": None, "unit": "cell*radius/s", "cores": args.cpu_threads, "kind": "port",
                   "sample": f"unavailable: {e}"}
    d.close()

    out = {
        "metric": METRIC,
        "value": value, "unit": "cell*radius/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPES[args.mode],
        "data": "synthetic",
        "config": {"workload": f"{N}^3 box {box:g} Mpc/h, {NS} radii 1 cell..{rmax:g} Mpc/h, "
                               f"zeta 40, Mmin 1e8 Msun/h, GRF seed 1", "mode": args.mode,
                   "radii_per_batch": nb, "x_pass_radii_per_launch": min(nb, fx_nb),
                   "fft_plans": ({"x": ("matrix" if plan["x_plan"] else "vector"),
                                  "z": ("matrix" if plan["z_plan"] else "vector")} if plan else None),
                   "autotuned": tuning is not None},
        "graph_ms_per_step": graph_ms,
        "graph_ms_per_step_median": float(np.median(step_ms)),
        "roofline": roof,
        "kernels_ms_last_step": {k: v[0] for k, v in kernels.items()},
        "plain_fp32": plain,
        "speedup_vs_plain_fp32": (value / plain["value"]) if plain else None,
        **({f"{k}_path": v for k, v in others.items()}),
        "cpu_baseline": cpu,
        **({"history_correction": hc} if hc else {}),
    }
    print(json.dumps(out))


def path_rate(lp, mode: str, steps: int, tune: bool = True) -> dict:
    """Another fused path on the same box (same generated field): autotuned
    like the main line, captured runs, warmed up once, wall time of `steps`
    runs."""
    from prc_d2z import D2Z
    N, NS = lp.N, lp.Nscales
    d = D2Z(lp.copy(mode=MODES[mode]))
    d.generate_delta(seed=1)
    if tune:
        d.autotune()
    d.run()
    d.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        d.run_async()
    d.sync()
    el = time.perf_counter() - t0
    plan = d.plan()
    d.close()
    return {"value": N ** 3 * NS * steps / el, "ms_per_step": el / steps * 1e3, "steps": steps,
            "dtype": DTYPES[mode], "fft_plans": {"x": "matrix" if plan["x_plan"] else "vector",
                                                  "z": "matrix" if plan["z_plan"] else "vector"}}


def plain_fp32_rate(lp, steps: int) -> dict:
    """The plain FP32 path (one radius per pass, no fusion) on the same box
    (same generated field): the comparison the metric names."""
    from prc_d2z import D2Z
    N, NS = lp.N, lp.Nscales
    d = D2Z(lp.copy(mode=MODES["plain"], radii_per_launch=0))
    d.generate_delta(seed=1)
    d.run()
    d.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        d.run_async()
    d.sync()
    el = time.perf_counter() - t0
    d.close()
    return {"value": N ** 3 * NS * steps / el, "ms_per_step": el / steps * 1e3, "steps": steps}


def multibox(args):
    """48 boxes of three sizes as one group (prc_d2z_group_run: value): the
    boxes spread over four chains (streams) by cells, each chain running its
    boxes of one size through one launch per pass, one graph per step;
    against the same boxes submitted together on their own streams
    (prc_d2z_run_many) and one after another as separate plain FP32 launches.
    value = aggregate cell-radius updates per second."""
    from prc_d2z import D2Z, D2ZParams, Group, run_many

    mode = MODES["fused"] if args.mode in ("mixed", "half") else MODES[args.mode]
    boxes, seed, cells = [], 1, 0
    for N, box in MULTIBOX:
        for _ in range(16):
            d = D2Z(D2ZParams(N=N, BoxSize=box, Nscales=48, Rmax=50.0, zeta=40.0, M_min=1.4706e8, mode=mode))
            d.generate_delta(seed)
            seed += 1
            boxes.append(d)
            cells += N ** 3
    grp = Group(boxes)
    for _ in range(max(1, args.warmup)):
        grp.run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        grp.run()
    batched = (time.perf_counter() - t0) / args.steps
    for _ in range(max(1, args.warmup)):
        run_many(boxes)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_many(boxes)
    streams = (time.perf_counter() - t0) / args.steps
    grp.close()
    # reference point: the same 48 boxes as separate plain launches, one after another
    seq_t = 0.0
    for k, b in enumerate(boxes):
        p = b.params.copy(mode=MODES["plain"])
        b.close()
        d = D2Z(p)
        d.generate_delta(k + 1)
        d.set_graph(False)
        d.run()
        t1 = time.perf_counter()
        d.run()
        seq_t += time.perf_counter() - t1
        d.close()
    value = cells * 48 / batched
    out = {
        "metric": "aggregate cell*radius updates/s, 48 uneven boxes (delta2zreion)",
        "value": value, "unit": "cell*radius/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": batched * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "16x128^3 (200 Mpc/h) + 16x192^3 (300) + 16x256^3 (400), 48 radii each "
                               "1 cell..50 Mpc/h, distinct GRF seeds",
                   "mode": "fused FP32 kernels; the boxes on four chains (streams) balanced by cells, "
                           "one launch per pass for the boxes of one size on a chain, one graph per step"},
        "streams_ms": streams * 1e3,
        "speedup_vs_streams": streams / batched,
        "separate_plain_ms": seq_t * 1e3,
        "speedup_vs_separate_plain": seq_t / batched,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
