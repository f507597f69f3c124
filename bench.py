#!/usr/bin/env python3
"""Headline benchmark: LaplCube Dirichlet solves/s at 128^3 (interior 127^3),
FP32+refine to 1e-10 next to plain FP64, plus the serving latency the metric
names and the other workloads as extra fields of the same JSON line.

Legs (each can be skipped):
  head  (the `value`): a step is one batched solve of `--batch` independent
        127^3 problems (fdmgpu_solve_batch_device), inputs resident in HBM.
        Inputs: seeded manufactured fields, u ~ U[-1,1) from a counter-based
        splitmix64 (seed 3 + problem index), f = the 7-point discrete Laplacian
        of u with zero Dirichlet boundary, geometry [-pi,pi]^3, h = 2pi/127,
        l = 128 h.
  c2    one 255^3 problem (seed 2): plain FP64 vs FP32+refine, error, iterations.
  c3    512 problems, each axis n in {31, 63, 127} by seed 3, 25% 2-D LaplRect,
        FP32+refine, kernel variants from the autotuner, all in one group
        (one launch per stage and solver, replayed as one hipGraph).
  c4    serving: Poisson arrivals (seeded exponential gaps) of the c3 mix plus
        5% 255^3, FP32+refine, at 0.5x / 0.8x / 1.0x of the measured
        saturation rate; p50/p99 latency from the scheduled arrival to the
        completion; FP64 one-at-a-time solves of the same mix beside it.
  c5    NSCube lid-driven cavity 127^3 (Re 250, dt 0.01, u0 1), 1000 steps
        with the FP32+refine projection; steps/s; compared with the CPU
        oracle's FP64 NSCube at step 100.
  pub   periodic LaplCube 32^3 / 64^3 / 128^3 solved one at a time (FP64
        and FP32) beside the reference's published n-body Poisson-stage
        timings (test/nbody_result.org).
  cpu   the CPU restatement of LaplCube<double> (oracle/fdm_oracle.cpp) on the
        host cores, OMP_PROC_BIND=close (bounded sample); the reference's own
        build beside it when present.

Prints ONE JSON line.
"""
import argparse
import json
import math
import os
import statistics
import subprocess
import sys
import time

# the CPU baseline's OpenMP threads stay on neighbouring cores (set before the
# oracle library, and with it the OpenMP runtime, is loaded)
os.environ.setdefault("OMP_PROC_BIND", "close")
# CPUs this process may run on (read before OpenMP binds the main thread to one)
CPUS_AT_START = len(os.sched_getaffinity(0))

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


# ---------------------------------------------------------------------------
# workload generation (numpy; identical definition to the test fixtures)
# ---------------------------------------------------------------------------
def splitmix_field(seed: int, n: int) -> np.ndarray:
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + (idx + np.uint64(1)) * np.uint64(0xD1B54A32D192ED03)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return 2.0 * (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 - 1.0


def lapl7(u: np.ndarray, h: float) -> np.ndarray:
    p = np.pad(u, 1)
    return (p[1:-1, 1:-1, :-2] + p[1:-1, 1:-1, 2:] + p[1:-1, :-2, 1:-1] + p[1:-1, 2:, 1:-1]
            + p[:-2, 1:-1, 1:-1] + p[2:, 1:-1, 1:-1] - 6.0 * u) / (h * h)


def manufactured(n: int, seed: int):
    h = 2 * math.pi / n
    u = splitmix_field(seed, n ** 3).reshape(n, n, n)
    return u, lapl7(u, h), h, (n + 1) * h


def lapl_aniso(u: np.ndarray, hs) -> np.ndarray:
    """7-point (3-D) or 5-point (2-D) discrete Laplacian, zero Dirichlet boundary,
    spacing hs[axis] per array axis."""
    p = np.pad(u, 1)
    out = np.zeros_like(u)
    for ax in range(u.ndim):
        lo = [slice(1, -1)] * u.ndim
        hi = [slice(1, -1)] * u.ndim
        lo[ax] = slice(0, -2)
        hi[ax] = slice(2, None)
        out += (p[tuple(lo)] + p[tuple(hi)] - 2.0 * u) / (hs[ax] * hs[ax])
    return out


def shape_geometry(kind: str, nx: int, ny: int, nz: int):
    """[-pi,pi] per axis: h = 2pi/n, l = (n+1) h (the NSCube convention)."""
    hx, hy, hz = 2 * math.pi / nx, 2 * math.pi / ny, 2 * math.pi / nz
    if kind == "rect":
        return (hx, hy, (nx + 1) * hx, (ny + 1) * hy)
    return (hx, hy, hz, (nx + 1) * hx, (ny + 1) * hy, (nz + 1) * hz)


def manufactured_shape(kind: str, nx: int, ny: int, nz: int, seed: int):
    if kind == "rect":
        u = splitmix_field(seed, nx * ny).reshape(ny, nx)
        return u, lapl_aniso(u, (2 * math.pi / ny, 2 * math.pi / nx))
    u = splitmix_field(seed, nx * ny * nz).reshape(nz, ny, nx)
    return u, lapl_aniso(u, (2 * math.pi / nz, 2 * math.pi / ny, 2 * math.pi / nx))


def uneven_mix(count: int, seed: int, big_frac: float = 0.0):
    """The batched-uneven mix: per problem each axis n from {31, 63, 127}, a 2-D
    LaplRect with probability 0.25; with probability big_frac a 255^3 cube."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        if big_frac and rng.random() < big_frac:
            out.append(("cube", 255, 255, 255))
        elif rng.random() < 0.25:
            nx, ny = (int(v) for v in rng.choice([31, 63, 127], 2))
            out.append(("rect", nx, ny, 1))
        else:
            nx, ny, nz = (int(v) for v in rng.choice([31, 63, 127], 3))
            out.append(("cube", nx, ny, nz))
    return out


def make_problem(serve, shp, d_ans, d_rhs, precision):
    kind, nx, ny, nz = shp
    g = shape_geometry(kind, nx, ny, nz)
    if kind == "rect":
        return serve.rect_problem(*g, nx, ny, d_ans, d_rhs, precision=precision)
    return serve.cube_problem(*g, nx, ny, nz, d_ans, d_rhs, precision=precision)


def progress(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pct(xs, q):
    return float(np.percentile(np.asarray(xs, np.float64), q)) if len(xs) else None


def unixbench_score(ts):
    """geometric mean of the fastest 2/3 (src/unixbench_score.cpp:9-20)."""
    ts = sorted(ts)
    keep = max(1, (2 * len(ts)) // 3)
    return math.exp(sum(math.log(t) for t in ts[:keep]) / keep)


# ---------------------------------------------------------------------------
# GPU legs
# ---------------------------------------------------------------------------
def algorithmic_bytes(kernel: str, mode: str, iters: int, n: int, batch: int, fused: bool = False) -> float:
    """Algorithmic HBM bytes per solve of `batch` problems for one kernel class.

    Per point: FP64 plain: xy_fwd 8+8, zcr 8+8, xy_inv 8+8 (48).
    FP32+refine with k inner solves (cubes whose planes fit LDS), default path:
      solve 1:  xy_fwd f64->f32 8+4, zcr 4+4, xy_inv f32->f32 4+4 (the
                iterate stays in FP32 scratch when a solve 2 follows; 4+8
                into the answer when k = 1);
      solve 2:  residual_rhs (read f, FP32 u; write r32) 8+4+4, xy_fwd_r32
                4+4, zcr 4+4, xy_inv_acc (read c, FP32 u; write u) 4+4+8;
      solve i >= 3: the same with the FP64 iterate: 8+8+4, 4+4, 4+4, 4+8+8;
      end:      residual, the norm-only certification (read f, u) 8+8.
    k = 2: 28 + 48 + 16 = 92 bytes/point.  Fused variant (fused=True): solve
    i >= 2 is xy_fwd_resid (residual inside the forward DST) 8+8+4, zcr 4+4,
    xy_inv_acc 4+8+8, and solve 1 writes the FP64 iterate: 96 at k = 2.
    """
    pts = float(n) ** 3 * batch
    if mode == "f64":
        per = {"xy_fwd": 16, "zcr": 16, "xy_inv": 16}.get(kernel, 0)
        return per * pts
    k = max(1, iters)
    if fused:
        first = {"xy_fwd": 12, "zcr": 8, "xy_inv": 12}.get(kernel, 0)
        later = {"xy_fwd_resid": 20, "zcr": 8, "xy_inv_acc": 20}.get(kernel, 0)
        return (first + later * (k - 1) + {"residual": 16}.get(kernel, 0)) * pts
    first = {"xy_fwd": 12, "zcr": 8, "xy_inv": 8 if k >= 2 else 12}.get(kernel, 0)
    second = {"residual_rhs": 16, "xy_fwd_r32": 8, "zcr": 8, "xy_inv_acc": 16}.get(kernel, 0) if k >= 2 else 0
    later = {"residual_rhs": 20, "xy_fwd_r32": 8, "zcr": 8, "xy_inv_acc": 20}.get(kernel, 0)
    return (first + second + later * max(0, k - 2) + {"residual": 16}.get(kernel, 0)) * pts


# Kernels of the 127^3 batch-32 workload as rocprofv3 names them: (key, label,
# symbol prefix, algorithmic bytes per point per launch).  Both FP32->FP32
# plane launches (solve 1's inverse into the FP32 iterate, solve 2's forward
# of the residual) are one symbol with the same bytes.
PROFILE_KERNELS = [
    ("xy_fwd", "forward planes, f64 -> f32 (solve 1)", "plane_dst_kernel<double, float, float, 7, 0, 0>", 12),
    ("plane_f32", "planes f32 -> f32 (solve 1 inverse, solve 2 forward)", "plane_dst_kernel<float, float, float, 7, 0, 0>", 8),
    ("zcr", "z cyclic reduction (FP32)", "zcr_strip_kernel<float, float, 7, 1, 16>", 8),
    ("residual_rhs", "FP64 residual -> FP32 rhs (solve 2)", "residual_tile_kernel<128, 8, true, float>", 16),
    ("xy_inv_acc", "inverse planes, u = u1 + c (solve 2)", "plane_dst_kernel<float, double, float, 7, 2, 0>", 16),
    ("residual", "FP64 residual, norm only", "residual_tile_kernel<128, 8, false, double>", 16),
    ("xy_f64", "planes FP64 (plain solve)", "plane_dst_kernel<double, double, double, 7, 0, 0>", 16),
    ("zcr_f64", "z cyclic reduction (FP64)", "zcr_strip_kernel<double, double, 7, 1, 16>", 16),
]
# timing class -> profile key, where they differ
PROFILE_KEY_OF_CLASS = {"xy_inv": "plane_f32", "xy_fwd_r32": "plane_f32"}


def profile_summary(kt: str, pmc: str, n: int = 127, batch: int = 32):
    """Writes profiles/lapl_cube_127_batch32_kernels.md and
    profiles/hbm_traffic_lapl_cube_127_batch32.json from the rocprofv3 csv
    output of the commands listed in the markdown (kernel trace in `kt`,
    FETCH_SIZE / WRITE_SIZE passes and their copy calibration in `pmc`)."""
    import csv
    import collections

    pts = float(n) ** 3 * batch

    def key_of(name):
        for key, _, sym, _ in PROFILE_KERNELS:
            if sym in name:
                return key
        return None

    trace = list(csv.DictReader(open(os.path.join(kt, "kt_kernel_trace.csv"))))
    stats = list(csv.DictReader(open(os.path.join(kt, "kt_kernel_stats.csv"))))
    dur = collections.defaultdict(list)
    for r in trace:
        k = key_of(r["Kernel_Name"])
        if k:
            dur[k].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)

    def counter(name, pre=""):
        path = os.path.join(pmc, pre + name, f"{pre}{name}_counter_collection.csv")
        agg = collections.defaultdict(list)
        for r in csv.DictReader(open(path)):
            k = "copy" if "copyBuffer" in r["Kernel_Name"] else key_of(r["Kernel_Name"])
            if k:
                agg[k].append(float(r["Counter_Value"]))
        return agg

    fetch, write = counter("FETCH_SIZE"), counter("WRITE_SIZE")
    cf, cw = counter("FETCH_SIZE", "copy_"), counter("WRITE_SIZE", "copy_")
    copy_kib = 256 * 1024
    kf = copy_kib / statistics.mean(cf["copy"])
    kw = copy_kib / statistics.mean(cw["copy"])
    out = {"n": n, "batch": batch, "fetch_scale": kf, "write_scale": kw, "kernels": {}}
    rows = []
    for key, label, sym, bpp in PROFILE_KERNELS:
        d = [x for x in dur.get(key, []) if x > 0.02]   # early-exit launches (converged problems) left out
        fl, wl = fetch.get(key, []), write.get(key, [])
        f = [x for x in fl if fl and x > 0.05 * max(fl)]
        w = [x for x in wl if wl and x > 0.05 * max(wl)]
        alg = bpp * pts
        rec = {"symbol": sym, "alg_bytes_per_launch": alg}
        if d:
            rec.update(avg_ms=statistics.mean(d), launches=len(d))
        if f and w:
            rec.update(fetch_kib=statistics.mean(f), write_kib=statistics.mean(w),
                       traffic_bytes_per_launch=round((statistics.mean(f) * kf + statistics.mean(w) * kw) * 1024))
        out["kernels"][key] = rec
        if d:
            gbs = alg / (rec["avg_ms"] / 1e3) / 1e9
            tr = rec.get("traffic_bytes_per_launch")
            rows.append(f"| {label} | `{sym}` | {len(d)} | {rec['avg_ms']:.4f} | {alg / 1e6:.1f} | {gbs:.0f} | "
                        f"{gbs / HBM_PEAK_GBS:.3f} | " + (f"{tr / 1e6:.1f} ({tr / alg:.2f}x)" if tr else "-") + " |")
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    with open(os.path.join(ROOT, "profiles", "hbm_traffic_lapl_cube_127_batch32.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    md = [
        "# LaplCube 127^3, batch 32: kernel trace and HBM traffic (MI355X)\n",
        "Workload: 32 independent Dirichlet 127^3 problems per step (N = 128), seeded manufactured fields, "
        "FP32+refine to 1e-10 (2 inner solves, default kernel variants) and plain FP64.\n",
        "Commands (from the repository root, one MI355X; the last one writes this file and the json beside it):\n",
        "```",
        f"rocprofv3 --kernel-trace --stats --output-format csv -d {kt} -o kt -- python3 bench.py --steps 5 --warmup 1 --no-cpu --legs head",
        f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d {pmc}/FETCH_SIZE -o FETCH_SIZE -- python3 bench.py --mode refine --steps 2 --warmup 1 --no-cpu",
        f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d {pmc}/WRITE_SIZE -o WRITE_SIZE -- python3 bench.py --mode refine --steps 2 --warmup 1 --no-cpu",
        f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d {pmc}/copy_FETCH_SIZE -o copy_FETCH_SIZE -- python3 bench.py --calib-copy-mib 256",
        f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d {pmc}/copy_WRITE_SIZE -o copy_WRITE_SIZE -- python3 bench.py --calib-copy-mib 256",
        f"python3 bench.py --profile-summary {kt} {pmc}",
        "```\n",
        f"Calibration: a 256 MiB device-to-device copy reads FETCH_SIZE = {copy_kib / kf:.0f} KiB and WRITE_SIZE = "
        f"{copy_kib / kw:.0f} KiB, so FETCH_SIZE is scaled by {kf:.2f} and WRITE_SIZE by {kw:.2f} to get bytes.\n",
        "Durations are per dispatch from the kernel trace.  Dispatches shorter than 20 us (launches for problems "
        "already converged, which read one flag and exit) are left out, and so are their counter readings.\n",
        "| kernel | symbol | dispatches | avg ms | algorithmic MB / launch | achieved GB/s | of 8 TB/s | "
        "HBM traffic MB / launch (x alg) |",
        "|---|---|---|---|---|---|---|---|",
    ] + rows + ["", "rocprofv3 --stats summary (all dispatches, warm-up included):", "",
                "| kernel | calls | avg ns | % |", "|---|---|---|---|"]
    for r in stats:
        md.append(f"| `{r['Name'].split('(')[0].replace('void ', '')}` | {r['Calls']} | {float(r['AverageNs']):.0f} | "
                  f"{float(r['Percentage']):.2f} |")
    with open(os.path.join(ROOT, "profiles", "lapl_cube_127_batch32_kernels.md"), "w") as fh:
        fh.write("\n".join(md) + "\n")
    print("\n".join(md))


def measured_traffic(kernel: str, n: int, batch: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    when it was taken on this workload; None otherwise."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_lapl_cube_127_batch32.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    if t.get("n") != n or t.get("batch") != batch:
        return None
    k = t.get("kernels", {}).get(PROFILE_KEY_OF_CLASS.get(kernel, kernel))
    return None if k is None else k.get("traffic_bytes_per_launch")


def run_gpu(mode: str, n: int, batch: int, steps: int, warmup: int, f_host: np.ndarray):
    import fdm_amd
    from fdm_amd import hip

    h = 2 * math.pi / n
    l = (n + 1) * h
    solver = fdm_amd.LaplCube(h, h, h, l, l, l, n, n, n, precision="refine" if mode == "refine" else "f64")
    d_f = hip.DeviceArray.from_host(f_host)
    d_u = hip.DeviceArray(f_host.shape, np.float64)
    stream = hip.Stream()
    max_iter = None
    for w in range(max(1, warmup)):
        solver.solve_batch(d_u, d_f, stream)
        if w == 0 and mode == "refine":
            # launch exactly the inner solves the first solve needed (no empty
            # launches for converged problems); the timed solves' residuals are
            # reported and checked against the tolerance
            max_iter = max(1, solver.stats().iterations)
            solver.set_refine(1e-10, max_iter)
    stream.synchronize()
    # per-kernel breakdown: a pass of its own with events around every launch
    bsteps = max(1, steps)
    solver.set_timing(True)
    for _ in range(bsteps):
        solver.solve_batch(d_u, d_f, stream)
    breakdown = {k: (v[0] / bsteps, v[1] / bsteps) for k, v in solver.timing().items()}
    dom = max((k for k in breakdown if k != "other"), key=lambda k: breakdown[k][0])
    # timed region: events around the dominant kernel's launches only (its
    # average launch time is the roofline figure; the other launches run
    # uninstrumented)
    solver.set_timing(True, kernels=[dom])
    ev0, ev1 = hip.Event(), hip.Event()
    hip.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        solver.solve_batch(d_u, d_f, stream)
    ev1.record(stream)
    stream.synchronize()
    hip.synchronize()
    t1 = time.perf_counter()
    wall_ms = (t1 - t0) * 1e3
    ev_ms = ev0.elapsed_ms(ev1)
    timing = solver.timing()
    st = solver.stats()
    u = d_u.download()
    iters = st.iterations if mode == "refine" else 1
    return {
        "wall_ms": wall_ms, "event_ms": ev_ms, "timing": timing, "breakdown": breakdown, "dominant": dom,
        "iterations": iters, "max_rel_residual": st.max_rel_residual, "u0": u[0], "u_last": u[-1],
        "max_iter": max_iter,
    }


def leg_host(n: int = 127, reps: int = 10):
    """The drop-in's host-array solve (fdmgpu_solve_host: pageable host arrays
    in and out, as LaplCube::solve(ans, rhs) receives them), PCIe copies
    included; one problem per call."""
    import fdm_amd
    u, f = manufactured_shape("cube", n, n, n, 5)
    g = shape_geometry("cube", n, n, n)
    out = {"workload": f"LaplCube Dirichlet {n}^3, host arrays (pageable) in and out, one solve per call"}
    for mode in ("f64", "refine"):
        s = fdm_amd.LaplCube(*g, n, n, n, precision=mode)
        if mode == "refine":
            s.set_refine(1e-10, 2)
        ans = np.zeros_like(f)
        s.solve(ans, f)
        t0 = time.perf_counter()
        for _ in range(reps):
            s.solve(ans, f)
        ms = (time.perf_counter() - t0) * 1e3 / reps
        out[mode] = {"ms_per_solve": round(ms, 3), "solves_per_s": round(1e3 / ms, 1),
                     "err_vs_manufactured": float(np.linalg.norm(ans - u) / np.linalg.norm(u)),
                     "pcie_bytes_per_solve": 2 * f.nbytes}
        del s
    return out


def leg_c2(reps: int = 5):
    """One 255^3 problem: plain FP64 vs FP32+refine (axis-tile kernels: a 256^2
    plane does not fit LDS)."""
    import fdm_amd
    from fdm_amd import hip
    n = 255
    u, f = manufactured_shape("cube", n, n, n, 2)
    g = shape_geometry("cube", n, n, n)
    d_f = hip.DeviceArray.from_host(f)
    out = {"workload": "LaplCube Dirichlet 255^3 (N=256), one problem, seed 2", "points": n ** 3}
    st = hip.Stream()
    res = {}
    for mode in ("f64", "refine"):
        s = fdm_amd.LaplCube(*g, n, n, n, precision=mode)
        if mode == "refine":
            # two inner solves reach 1e-10 here (the result reports the residual
            # and whether it converged); each launched but unneeded iteration
            # costs seven no-op kernels
            s.set_refine(1e-10, 2)
        d_u = hip.DeviceArray(f.shape, np.float64)
        s.solve(d_u, d_f, st)
        st.synchronize()
        e0, e1 = hip.Event(), hip.Event()
        e0.record(st)
        for _ in range(reps):
            s.solve(d_u, d_f, st)
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        # the kernel breakdown in a pass of its own: per-launch events add
        # about a fifth to a single solve of this size
        s.set_timing(True)
        for _ in range(reps):
            s.solve(d_u, d_f, st)
        st.synchronize()
        stats = s.stats()
        ans = d_u.download()
        it = stats.iterations if mode == "refine" else 1
        res[mode] = ans
        out[mode] = {"ms_per_solve": round(ms, 4), "solves_per_s": round(1e3 / ms, 2), "iterations": it,
                     "err_vs_manufactured": float(np.linalg.norm(ans - u) / np.linalg.norm(u)),
                     "model_GBs": round(48.0 * it * n ** 3 / (ms / 1e3) / 1e9, 1),
                     "kernel_ms_per_solve": {k: round(v[0] / reps, 4) for k, v in s.timing().items()}}
        if mode == "refine":
            out[mode]["rel_residual"] = stats.rel_residual
            out[mode]["max_iter"] = 2
            out[mode]["converged"] = bool(stats.rel_residual <= 1e-10)
        del s
    out["err_refine_vs_fp64_gpu"] = float(np.linalg.norm(res["refine"] - res["f64"]) / np.linalg.norm(res["f64"]))
    out["refine_over_fp64"] = round(out["refine"]["solves_per_s"] / out["f64"]["solves_per_s"], 3)
    return out


def leg_c3(steps: int = 5, count: int = 512, tune_reps: int = 3):
    """512 uneven problems, FP32+refine, autotuned variants, one group replayed
    as a hipGraph."""
    import fdm_amd
    from fdm_amd import hip, serve
    mix = uneven_mix(count, 3)
    t0 = time.perf_counter()
    bufs, probs, us = [], [], []
    points = 0
    for i, shp in enumerate(mix):
        u, f = manufactured_shape(*shp, 1000 + i)
        d_f = hip.DeviceArray.from_host(f)
        d_u = hip.DeviceArray(f.shape, np.float64)
        bufs.append((d_u, d_f))
        probs.append(make_problem(serve, shp, d_u, d_f, fdm_amd.F32_REFINE64))
        if i < 8:
            us.append(u)
        points += f.size
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    serve.tune(probs, reps=tune_reps)
    tune_s = time.perf_counter() - t0
    grp = serve.Group(probs, nstreams=4, graph=True)
    st = hip.Stream()
    grp.solve(st)          # eager, up to 4 inner solves
    # launch exactly the inner solves the eager solve needed (no empty launches
    # for converged problems), as the head leg does; every timed result's
    # residual is checked below
    max_iter = max(grp.result(i).iterations for i in range(count))
    grp.set_refine(1e-10, max_iter)
    grp.solve(st)          # eager again with the new limit
    grp.solve(st)          # captured, first replay
    st.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(st)
    for _ in range(steps):
        grp.solve(st)
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    res = [grp.result(i) for i in range(count)]
    errs = [float(np.linalg.norm(bufs[i][0].download() - us[i]) / np.linalg.norm(us[i])) for i in range(len(us))]
    iters = [r.iterations for r in res]
    out = {
        "workload": f"{count} problems, each axis n in {{31,63,127}} by seed 3, 25% 2-D LaplRect, FP32+refine "
                    f"to 1e-10; one group (one launch per stage and solver, 4 streams) replayed as a hipGraph",
        "problems": count, "solvers": len({m for m in mix}), "points": points,
        "ms_per_step": round(ms, 3), "solves_per_s": round(count / (ms / 1e3), 1),
        "points_per_s": round(points / (ms / 1e3), 1),
        "model_GBs": round(sum(48.0 * it * float(np.prod([d for d in m[1:]])) for it, m in zip(iters, mix))
                           / (ms / 1e3) / 1e9, 1),
        "iterations_max": max(iters), "max_iter": max_iter, "max_rel_residual": max(r.rel_residual for r in res),
        "all_converged": all(r.status == 0 for r in res), "err_vs_manufactured_first8": max(errs),
        "tuned": True, "tune_s": round(tune_s, 2), "gen_s": round(gen_s, 2),
    }
    grp.close()
    return out


def leg_c4(duration_s: float = 2.0, budget_gb: float = 24.0, max_batch: int = 64, seed: int = 4,
           fp64_sample: int = 120, fracs=(0.25, 0.5, 0.75), rates=None, sat_requests: int = 1200):
    """Serving: Poisson arrivals of the uneven mix + 5% 255^3 (FP32+refine).
    Saturation is the throughput of a burst of `sat_requests` requests (one
    255^3 request in twenty carries about three quarters of the mix's points,
    so the burst must be long for its share to be typical); the Poisson runs
    offer `fracs` of it, or the absolute `rates` (requests/s) when given."""
    import fdm_amd
    from fdm_amd import hip, serve
    rng = np.random.default_rng(seed)
    depth = 24                   # in-flight requests per shape (own rhs and ans buffers each)
    # request streams first, so every buffer exists before anything is timed
    warm = uneven_mix(200 + sat_requests, seed, big_frac=0.05)
    fp64_reqs = uneven_mix(fp64_sample, seed + 7, big_frac=0.05)
    shapes = {}

    def prepare(shps):
        for shp in shps:
            if shp in shapes:
                continue
            u, f = manufactured_shape(*shp, 2000 + len(shapes))
            tmpl = hip.DeviceArray.from_host(f)
            slots = []
            for _ in range(depth):
                d_f = hip.DeviceArray(f.shape, np.float64)
                hip.check(hip.rt().hipMemcpy(d_f.ptr, tmpl.ptr, d_f.nbytes, hip.D2D), "copy")
                slots.append((hip.DeviceArray(f.shape, np.float64), d_f))
            shapes[shp] = {"f": f, "tmpl": tmpl, "slots": slots, "busy": [None] * depth, "next": 0}

    prepare(warm + fp64_reqs)
    S = serve.Server(hbm_budget=int(budget_gb * (1 << 30)), max_batch=max_batch, nstreams=4)

    def submit(shp, t_sched):
        e = shapes[shp]
        j = e["next"]
        e["next"] = (j + 1) % depth
        if e["busy"][j] is not None:      # slot reuse: its previous request must be done
            S.wait(e["busy"][j])
        d_u, d_f = e["slots"][j]
        t_sub = time.perf_counter()
        tk = S.submit(make_problem(serve, shp, d_u, d_f, fdm_amd.F32_REFINE64))
        e["busy"][j] = tk
        return tk, t_sub - t_sched

    def run(reqs, gaps):
        """Submit at the scheduled times; latency = client delay + server latency."""
        prepare(reqs)
        t_start = time.perf_counter()
        sched = t_start + np.concatenate([[0.0], np.cumsum(gaps)[:-1]]) if gaps is not None else None
        tick = []
        for i, shp in enumerate(reqs):
            ts = sched[i] if sched is not None else time.perf_counter()
            while time.perf_counter() < ts:
                pass
            tick.append(submit(shp, ts))
        t_subs = time.perf_counter()
        S.drain()
        t_end = time.perf_counter()
        run.submit_s = t_subs - t_start
        lat, resid, ok = [], [], True
        for tk, delay in tick:
            r = S.poll(tk)
            lat.append(r.latency_ms + 1e3 * delay)
            resid.append(r.rel_residual)
            ok = ok and r.status == 0
            S.forget(tk)
        for e in shapes.values():
            e["busy"] = [None] * depth
        return t_end - t_start, lat, resid, ok

    # warm-up (plans, tables, pinned batches, iteration counts), then saturation:
    # a burst of requests all submitted at t = 0
    run(warm[:200], None)
    dt, lat, _, _ = run(warm[200:], None)
    sat = sat_requests / dt
    burst_submit_per_s = sat_requests / run.submit_s
    out = {"workload": "requests of the uneven mix (each axis n in {31,63,127}, 25% 2-D) + 5% 255^3, "
                       "FP32+refine to 1e-10, Poisson arrivals (seeded exponential gaps)",
           "hbm_budget_gb": budget_gb, "max_batch": max_batch, "streams": 4,
           "saturation_solves_per_s": round(sat, 1), "burst_client_submit_per_s": round(burst_submit_per_s, 1),
           "rates": []}
    offered = [(r, 1000 + k) for k, r in enumerate(rates)] if rates is not None else \
        [(frac * sat, int(frac * 100)) for frac in fracs]
    for rate, tag in offered:
        n = max(50, int(rate * duration_s))
        reqs = uneven_mix(n, seed + tag, big_frac=0.05)
        gaps = rng.exponential(1.0 / rate, n)
        dt, lat, resid, ok = run(reqs, gaps)
        out["rates"].append({"offered_per_s": round(rate, 1), "requests": n,
                             "achieved_solves_per_s": round(n / dt, 1),
                             "p50_ms": round(pct(lat, 50), 3), "p99_ms": round(pct(lat, 99), 3),
                             "max_rel_residual": max(resid), "all_converged": ok})
    S.close()
    # FP64 one request at a time (no batching, no overlap), same mix
    plans, times = {}, []
    st = hip.Stream()
    for shp in fp64_reqs:
        d_u, d_f = shapes[shp]["slots"][0]
        if shp not in plans:
            kind, nx, ny, nz = shp
            g = shape_geometry(*shp)
            plans[shp] = (fdm_amd.LaplRect(*g, nx, ny, precision="f64") if kind == "rect"
                          else fdm_amd.LaplCube(*g, nx, ny, nz, precision="f64"))
            plans[shp].solve(d_u, d_f, st)   # warm-up
            st.synchronize()
        t = time.perf_counter()
        plans[shp].solve(d_u, d_f, st)
        st.synchronize()
        times.append(1e3 * (time.perf_counter() - t))
    out["fp64_one_at_a_time"] = {"requests": len(times), "solves_per_s": round(1e3 * len(times) / sum(times), 1),
                                 "p50_ms": round(pct(times, 50), 3), "p99_ms": round(pct(times, 99), 3)}
    return out


# The reference's only published timings for this path: its n-body code's
# "Poisson" stage, one periodic LaplCube solve (+ one rhs scaling sweep), in ms
# (test/nbody_result.org:9,24,39): Apple M1 CPU / M1 GPU (Vulkan) / Ryzen 7
# 3700X (16 OpenMP threads) / RTX 3060 (Vulkan).
PUBLISHED_POISSON_MS = {32: (1.7, 1.21, 1.26, 0.25), 64: (8.87, 5.69, 6.90, 0.86), 128: (68.83, 40.37, 62.50, 5.50)}


def leg_published(reps: int = 50):
    """Periodic LaplCube N^3 (N = 32, 64, 128: all N points per axis), one
    solve at a time on a stream, device-resident, FP64 and FP32: the
    reference's published Poisson-stage configuration.  The reference's timed
    window also holds one rhs scaling sweep, not timed here."""
    import fdm_amd
    from fdm_amd import hip
    out = {"workload": "periodic LaplCube N^3 (tensor_flags<periodic,periodic,periodic>), one solve at a time, "
                       "inputs resident in HBM; published: the reference's n-body Poisson stage",
           "published_ms": {str(k): dict(zip(("m1_cpu", "m1_gpu_vulkan", "ryzen_3700x_16t", "rtx3060_vulkan"), v))
                            for k, v in PUBLISHED_POISSON_MS.items()}}
    st = hip.Stream()
    for n in (32, 64, 128):
        h = 2 * math.pi / n
        rhs = np.random.default_rng(n).uniform(-1, 1, (n, n, n))
        rhs -= rhs.mean()
        row = {}
        for prec, dt in (("f64", np.float64), ("f32", np.float32)):
            s = fdm_amd.LaplCube(h, h, h, 2 * math.pi, 2 * math.pi, 2 * math.pi, n, n, n, precision=prec,
                                 periodic=True)
            d_f = hip.DeviceArray.from_host(rhs.astype(dt))
            d_u = hip.DeviceArray(rhs.shape, dt)
            for _ in range(3):
                s.solve(d_u, d_f, st)
            st.synchronize()
            e0, e1 = hip.Event(), hip.Event()
            e0.record(st)
            for _ in range(reps):
                s.solve(d_u, d_f, st)
            e1.record(st)
            e1.synchronize()
            ms = e0.elapsed_ms(e1) / reps
            row[prec] = {"ms_per_solve": round(ms, 4), "solves_per_s": round(1e3 / ms, 1)}
            del s
        row["speedup_vs_rtx3060_f64"] = round(PUBLISHED_POISSON_MS[n][3] / row["f64"]["ms_per_solve"], 1)
        row["speedup_vs_ryzen_f64"] = round(PUBLISHED_POISSON_MS[n][2] / row["f64"]["ms_per_solve"], 1)
        out[str(n)] = row
    return out


def leg_c5(n: int = 127, steps: int = 1000, check_at: int = 100, cpu_check: bool = True):
    """NSCube lid-driven cavity n^3 on [-pi,pi]^3, Re 250, dt 0.01, u0 1, from
    rest; pressure projection = the FP32+refine solve.  Compared with the CPU
    oracle's FP64 NSCube at step `check_at`, then timed to `steps`."""
    import fdm_amd
    from fdm_amd import hip
    import oracle as O
    kw = dict(u0=1.0, Re=250.0, dt=0.01)
    g = fdm_amd.NSCube(n, n, n, solver="refine", **kw)
    out = {"workload": f"NSCube lid-driven cavity {n}^3 on [-pi,pi]^3, Re 250, dt 0.01, u0 1, from rest; "
                       "pressure solve FP32+refine to 1e-10", "steps": steps}
    st = hip.Stream()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(st)
    g.step(check_at, st)
    e1.record(st)
    e1.synchronize()
    first_ms = e0.elapsed_ms(e1)
    if cpu_check:
        o = O.NsCube(n, n, n, **kw)
        t = time.perf_counter()
        o.step(check_at)
        cpu_s = time.perf_counter() - t
        out["compare_at_step"] = check_at
        out["max_rel_diff_vs_cpu_fp64"] = {f: float(np.abs(g.get(f) - o.get(f)).max() / np.abs(o.get(f)).max())
                                           for f in ("u", "v", "w", "p")}
        out["cpu_oracle_steps_per_s"] = round(check_at / cpu_s, 2)
        del o
    rest = steps - check_at
    e0.record(st)
    g.step(rest, st)
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_ms(e1)
    s = g.solver_stats()
    u = g.get("u")
    out.update({"steps_per_s": round(rest / (ms / 1e3), 1), "ms_per_step": round(ms / rest, 4),
                "first_steps_ms_per_step": round(first_ms / check_at, 4), "last_solve_rel_residual": s.rel_residual,
                "last_solve_iterations": s.iterations, "u_max": float(np.abs(u[1:n + 1]).max()),
                "finite": bool(np.isfinite(u).all())})
    g.close()
    # beside it: the same run with the plain FP64 pressure solve (what the
    # drop-in LaplCube<double> uses by default), and its distance to the
    # refined run after the same number of steps
    g64 = fdm_amd.NSCube(n, n, n, solver="f64", **kw)
    g64.step(check_at, st)
    st.synchronize()
    e0.record(st)
    g64.step(rest, st)
    e1.record(st)
    e1.synchronize()
    ms64 = e0.elapsed_ms(e1)
    u64 = g64.get("u")
    out["fp64_solver"] = {"steps_per_s": round(rest / (ms64 / 1e3), 1), "ms_per_step": round(ms64 / rest, 4),
                          "u_max_rel_diff_vs_refine": float(np.abs(u64 - u).max() / np.abs(u).max())}
    g64.close()
    return out


def host_cpu_info():
    """CPU model, physical cores and the logical CPUs this process may run on."""
    model, sockets, cores_per_socket = None, 1, None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            key, val = key.strip(), val.strip()
            if key == "Model name":
                model = val
            elif key == "Socket(s)":
                sockets = int(val)
            elif key == "Core(s) per socket":
                cores_per_socket = int(val)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return {"cpu_model": model, "physical_cores": (sockets * cores_per_socket) if cores_per_socket else None,
            "cpus_available": CPUS_AT_START}


def time_solves(solve, budget_s: float):
    """1 warm-up, then >= 20 timed solves (up to budget_s, at most 200)."""
    solve()
    times = []
    t_start = time.perf_counter()
    while len(times) < 20 or (time.perf_counter() - t_start < budget_s and len(times) < 200):
        t = time.perf_counter()
        solve()
        times.append(time.perf_counter() - t)
        if time.perf_counter() - t_start > 4 * budget_s:
            break
    return times


def cpu_baseline(n: int, f0: np.ndarray, budget_s: float = 15.0):
    """The CPU restatement of LaplCube<double,false> (oracle/fdm_oracle.cpp, the
    reference's pass order, factors and transforms), OpenMP over this process's
    host cores with OMP_PROC_BIND=close; when the reference's own build
    (oracle/_ref) is present, a short sample of it is timed beside it."""
    import oracle as O
    h = 2 * math.pi / n
    l = (n + 1) * h
    threads = O.lib().orc_num_threads()
    ans = None

    def port():
        nonlocal ans
        ans = O.lapl_cube(f0, h, h, h, l, l, l)

    times = time_solves(port, budget_s)
    score = unixbench_score(times)
    med = statistics.median(times)
    info = host_cpu_info()
    out = {
        "value": 1.0 / score, "unit": "solves/s", "cores": threads, "kind": "port",
        "sample": f"{len(times)} LaplCube<double> Dirichlet {n}^3 solves of the CPU restatement after 1 warm-up, "
                  f"one at a time, {threads} OpenMP threads (OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}) on "
                  f"{info['cpu_model']} ({info['physical_cores']} physical cores, {info['cpus_available']} CPUs "
                  f"available to the process); value = 1/unixbench_score (geomean of fastest 2/3); median "
                  f"{1.0 / med:.2f} solves/s",
        "median_solves_per_s": 1.0 / med,
        "host": info, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
    }
    if O.ref_available():   # the reference's own LaplCube<double> on the same cores
        R = O.ref()
        hd = R.ref_cube_create_d(h, h, h, l, l, l, n, n, n)
        ra = np.zeros_like(f0)

        def ref():
            R.ref_cube_solve_d(hd, ra, f0.copy())

        rt = time_solves(ref, min(5.0, budget_s))
        R.ref_cube_destroy(hd)
        out["reference_solves_per_s"] = 1.0 / unixbench_score(rt)
        out["reference_sample"] = f"{len(rt)} solves of the reference's own build (oracle/_ref), same cores"
    return out, ans


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=32, help="127^3 problems per step")
    ap.add_argument("--n", type=int, default=127)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--mode", choices=("both", "refine", "f64"), default="both",
                    help="which GPU legs to run (profiling runs use one)")
    ap.add_argument("--legs", default="head,c2,c3,c4,c5,pub,host,cpu",
                    help="comma list of legs: head (the value; always run), c2, c3, c4, c5, pub, host, cpu")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="arrival window per serving rate")
    ap.add_argument("--profile-summary", nargs=2, metavar=("KT_DIR", "PMC_DIR"),
                    help="summarise rocprofv3 csv output into profiles/ and exit (no GPU)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="only run device-to-device copies of this many MiB (calibrates the HBM "
                         "FETCH_SIZE/WRITE_SIZE counters against a copy of known size) and exit")
    args = ap.parse_args()
    if args.profile_summary:
        profile_summary(*args.profile_summary)
        return
    if args.calib_copy_mib:
        from fdm_amd import hip
        nb = args.calib_copy_mib << 20
        a = hip.DeviceArray((nb // 8,), np.float64)
        b = hip.DeviceArray((nb // 8,), np.float64)
        for _ in range(3):
            hip.check(hip.rt().hipMemcpy(b.ptr, a.ptr, nb, hip.D2D), "copy")
        hip.synchronize()
        print(json.dumps({"calib_copy_bytes": nb, "copies": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (multi-GPU sharding "
                                   "by whole problems is a design note, not implemented)"}))
        sys.exit(2)
    legs = set(args.legs.split(","))
    if args.no_cpu:
        legs.discard("cpu")
    n, B = args.n, args.batch

    # inputs: B seeded manufactured problems (seed 3 + i), as in the batched workload
    us, fs = [], []
    for i in range(B):
        u, f, _, _ = manufactured(n, 3 + i)
        fs.append(f)
        if i in (0, B - 1):
            us.append(u)
    F = np.stack(fs)
    del fs

    if args.mode != "both":
        r = run_gpu(args.mode, n, B, args.steps, args.warmup, F)
        print(json.dumps({"mode": args.mode, "solves_per_s": B * args.steps / (r["wall_ms"] / 1e3),
                          "iterations": r["iterations"],
                          "kernel_ms_per_step": {k: v[0] for k, v in r["breakdown"].items()}}))
        return
    progress(f"head: {B} x {n}^3 FP32+refine and FP64")
    res_r = run_gpu("refine", n, B, args.steps, args.warmup, F)
    res_d = run_gpu("f64", n, B, args.steps, args.warmup, F)
    fcpu = F[0].copy()

    solves = B * args.steps
    value = solves / (res_r["wall_ms"] / 1e3)
    fp64_value = solves / (res_d["wall_ms"] / 1e3)

    # dominant kernel of the headline (refine) path (from the breakdown pass)
    # and its roofline (its launches in the timed region)
    tm = res_r["breakdown"]
    dom = res_r["dominant"]
    ms_total, launches = res_r["timing"][dom]
    fused = tm.get("xy_fwd_resid", (0.0, 0))[1] > 0
    alg = algorithmic_bytes(dom, "refine", res_r["iterations"], n, B, fused) * args.steps
    achieved = alg / (ms_total / 1e3) / 1e9
    # whole-solve algorithmic bytes: the per-class model summed over the classes that ran
    model_bytes = sum(algorithmic_bytes(k, "refine", res_r["iterations"], n, B, fused) for k in tm) * args.steps
    err_u = float(np.linalg.norm(res_d["u0"] - us[0]) / np.linalg.norm(us[0]))
    err_r = float(np.linalg.norm(res_r["u0"] - res_d["u0"]) / np.linalg.norm(res_d["u0"]))

    out = {
        "metric": "solves/s @128\u00b3 to 1e-10 rel (FP32+refine vs FP64); p50/p99 serve latency",
        "value": round(value, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res_r["wall_ms"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 solve + fp64 residual/update",
        "data": "synthetic (seeded splitmix64 manufactured fields)",
        "config": {"workload": f"LaplCube Dirichlet {n}^3 interior (N={n + 1}) FP32+refine to 1e-10, "
                               f"batch of {B} per step", "n": n, "batch": B, "tolerance": 1e-10},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dom, n, B),
                     "alg_bytes_per_launch": round(alg / launches),
                     "avg_launch_ms": round(ms_total / launches, 5), "launches": launches},
        "iterations": res_r["iterations"],
        "max_iter": res_r["max_iter"],
        "max_rel_residual": res_r["max_rel_residual"],
        "converged": bool(0 <= res_r["max_rel_residual"] <= 1e-10),
        "fp64_solves_per_s": round(fp64_value, 2),
        "refine_over_fp64": round(value / fp64_value, 3),
        "err_fp64_vs_manufactured": err_u,
        "err_refine_vs_fp64_gpu": err_r,
        "solve_model_GBs": round(model_bytes / (res_r["wall_ms"] / 1e3) / 1e9, 1),
        "kernel_ms_per_step": {k: round(v[0], 4) for k, v in tm.items()},
        "fp64_kernel_ms_per_step": {k: round(v[0], 4) for k, v in res_d["breakdown"].items()},
    }
    del F
    if "c2" in legs:
        progress("c2: 255^3 FP64 vs FP32+refine")
        out["single_255"] = leg_c2()
    if "c3" in legs:
        progress("c3: 512 uneven problems (tune, group, graph)")
        out["batched_uneven"] = leg_c3()
    if "c4" in legs:
        progress("c4: serving at 3 arrival rates")
        out["serve"] = leg_c4(duration_s=args.serve_seconds)
        out["p50_ms"] = [r["p50_ms"] for r in out["serve"]["rates"]]
        out["p99_ms"] = [r["p99_ms"] for r in out["serve"]["rates"]]
    if "host" in legs:
        progress("host: drop-in host-array solves, PCIe copies included")
        out["host_solve"] = leg_host()
    if "pub" in legs:
        progress("pub: periodic LaplCube 32^3..128^3 one at a time (the reference's published Poisson stage)")
        out["published_poisson"] = leg_published()
    if "c5" in legs:
        progress("c5: NSCube 127^3 cavity, 1000 steps (CPU oracle check at step 100)")
        out["nscube"] = leg_c5(cpu_check="cpu" in legs)
    if "cpu" in legs:
        progress("cpu: CPU restatement of LaplCube<double> on the host cores")
        cb, ans_cpu = cpu_baseline(n, fcpu, args.cpu_budget)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["cpu_host"] = cb["host"]
        if "reference_solves_per_s" in cb:
            out["cpu_reference_solves_per_s"] = round(cb["reference_solves_per_s"], 2)
        out["err_fp64_gpu_vs_cpu"] = float(np.linalg.norm(res_d["u0"] - ans_cpu) / np.linalg.norm(ans_cpu))
        out["err_refine_vs_cpu"] = float(np.linalg.norm(res_r["u0"] - ans_cpu) / np.linalg.norm(ans_cpu))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
