"""Benchmark: Euler RK4 step throughput on one MI355X.

Workload (default): one curvilinear 4096x4096 block, upwind9 with upwind
dissipation, isentropic vortex (centre (-1, 0), r* 0.5, eps 1, Mach 0.5) plus
seeded +-1e-3 density noise, periodic SATs, dt = 0.2/4095, FP64. A step is one
classical RK4 step (four fused stage kernels) over the whole block with the
state resident in HBM.

Prints one JSON line: throughput in Gpoint-RK4-steps/s, the HBM roofline of
the stage kernel (algorithmic bytes per launch / mean launch duration from HIP
events), and the CPU oracle timed on this host on a bounded sample.
`--precision mixed` runs the FP32-RHS/FP64-state form; `--workload c3|c5|interp`
prints the same line for the 24-block Grid layout, the 256-solve batch and a
two-block layout with 2:1 interpolated interfaces.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "Gpoint-RK4-steps/s + % HBM roofline; mixed vs FP64 at ≤1e-8 rel L2 err"
UNIT = "Gpoint-RK4-steps/s"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per point per RK4 step of the fused schedule (each
# stage reads its input and the metric planes it uses once, plus the
# own-point operands of its stage combination, and writes its outputs once).
# FP64 (J is formed from the four metric products, so 4 metric planes):
# stage 1 reads y_n, metrics, writes Y2 (96 B); stage 2 reads Y2, metrics,
# y_n, writes Y3 (128 B); stage 3 reads Y3, metrics, y_n, Y2, writes Y4 and
# W = Y2 + 2 Y3 - y_n (192 B); stage 4 reads Y4, metrics, W and writes
# y_{n+1} (128 B): 544 B.
# The reference's schedule (running accumulator, 5 metric planes) needs
# 672 B; bench also reports the rate at which that schedule's bytes would
# have to move to match this kernel ("reference_schedule").
# Mixed: the stage inputs are y_n + Δ_s with FP32 increments Δ_s, and the
# last stage sums y_n and the increments in FP64: stage 1 reads y_n, metrics,
# writes Δ_1 (80 B); stages 2-3 read y_n, Δ_s, metrics and write Δ_{s+1}
# (96 B each); stage 4 reads y_n, Δ_3, metrics, Δ_1, Δ_2 and writes y_{n+1}
# (144 B): 416 B.
BYTES_PER_POINT_STEP = {"fp64": 544.0, "mixed": 416.0}
# advance_adaptive's embedded Bogacki-Shampine step, unfused: four RHS passes
# (stage input, 4 metric planes in, k out: 96 B each), the stage inputs
# (y_n + k0; y_n + k1; y_n + k0..k2: 96, 96, 160 B), fut (160 B), the
# embedded solution (192 B), the H-norm error (64 B) and the copy of an
# accepted step (64 B): 1216 B per point per accepted step.
BYTES_PER_POINT_ADAPTIVE_STEP = 1216.0
REFERENCE_SCHEDULE_BYTES = 672.0
# HBM bytes per stage launch from rocprofv3 counters (bench.py --counters)
TRAFFIC_FILES = {"fp64": os.path.join(ROOT, "profiles", "stage_kernel_upwind9_4096_traffic.json"),
                 "mixed": os.path.join(ROOT, "profiles", "stage_kernel_upwind9_4096_mixed_traffic.json")}


def _counter_rows(path, counter):
    """(dispatch id, kernel name, value) of one counter in a rocprofv3 --pmc
    CSV (run_counter_collection.csv), in dispatch order."""
    import csv
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return rows


def summarize_counters(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, n, tile_index, precision):
    """HBM bytes per stage launch from FETCH_SIZE / WRITE_SIZE (KiB per
    dispatch) of a run of this bench, scaled by what the same counters read
    for the 2 GiB device copy of `--calibrate-copy`. A stage launch is the
    interior-tile dispatch plus the edge-tile dispatch (consecutive in the
    trace); launches cycle through the four RK4 stages; the first step is
    left out (cold caches)."""
    copy_bytes = float(1 << 31)

    def factor(path, counter):
        vals = sorted(v for _, k, v in _counter_rows(path, counter) if "copyBuffer" in k)
        return copy_bytes / (vals[len(vals) // 2] * 1024.0)

    ff, wf = factor(copy_fetch_csv, "FETCH_SIZE"), factor(copy_write_csv, "WRITE_SIZE")

    def per_launch(path, counter, scale):
        rows = [(k, v) for _, k, v in _counter_rows(path, counter) if "stage_kernel" in k]
        if len(rows) % 2:
            raise SystemExit(f"{path}: odd number of stage dispatches")
        kernel = next(k for k, _ in rows if k.rstrip(")").split("(")[0].endswith("false>"))
        pairs = [(rows[i][1] + rows[i + 1][1]) * 1024.0 * scale for i in range(0, len(rows), 2)]
        stages = [[] for _ in range(4)]
        for i, v in enumerate(pairs[4:]):
            stages[i % 4].append(v)
        return kernel, [sum(s) / len(s) for s in stages]

    kernel, fetch = per_launch(fetch_csv, "FETCH_SIZE", ff)
    _, write = per_launch(write_csv, "WRITE_SIZE", wf)
    return {"tile_index": tile_index, "tile": list(_tile_of(tile_index)), "kernel": kernel.split("(")[0],
            "hbm_bytes_per_launch": (sum(fetch) + sum(write)) / 4.0,
            "algorithmic_bytes_per_launch": BYTES_PER_POINT_STEP[precision] / 4.0 * n * n,
            "fetch_factor": ff, "write_factor": wf,
            "per_stage": [{"stage": s, "fetch_bytes": fetch[s], "write_bytes": write[s]} for s in range(4)]}


def _tile_of(index):
    """(TX, TY, RY) of a tile shape index (a host-side table lookup)."""
    import ctypes
    from summationbyparts_amd import _lib
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _lib.check(_lib.lib().sbp_tile_shape_get(index, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
               "sbp_tile_shape_get")
    return a.value, b.value, c.value


class Workload:
    """A system ready to step, its step size, point count, description and a
    bounded CPU-oracle sample."""

    def __init__(self, sys_, dt, points, desc, cfg, cpu):
        self.sys, self.dt, self.points, self.desc, self.cfg, self.cpu = sys_, dt, points, desc, cfg, cpu


def _baseline(points, sec, cores, sample):
    return {"value": points / sec / 1e9, "unit": UNIT, "cores": cores, "kind": "port", "sample": sample}


def build_c2(n, precision, cpu_steps):
    """One curvilinear n x n block, upwind9 + dissipation, vortex + 1e-3
    density noise, periodic SATs, dt = 0.2/(n-1)."""
    from common import curvilinear, density_noise
    from summationbyparts_amd import _lib, euler, operators

    X, Y = curvilinear(n, n)
    sys_ = euler.System(X, Y, operators.Upwind9())
    if precision == "mixed":
        sys_.set_precision(_lib.MIXED)
    sys_.vortex(0.0, euler.VortexParameters([euler.Vortice(-1.0, 0.0, 0.5, 1.0)], 0.5))
    q = sys_.field()
    q[0] += density_noise(1, 0, n, n)
    sys_.set_state(0, q)
    sys_.set_rhs_kind(_lib.RHS_UPWIND)

    def cpu():
        # the oracle, one thread: the reference's one-thread-per-block semantics for one block
        from oracle import oracle as O
        ref = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
        ref.set_state(0, q)
        t0 = time.perf_counter()
        ref.advance(0.2 / (n - 1), cpu_steps, 1)
        sec = time.perf_counter() - t0
        return _baseline(n * n * cpu_steps, sec, 1,
                         f"{cpu_steps} RK4 steps of the {n}x{n} upwind9 curvilinear block, oracle (C++ restatement), "
                         f"1 thread, {os.cpu_count()} host cores visible")

    desc = (f"single curvilinear block {n}x{n}, upwind9 + dissipation, vortex + 1e-3 density noise, periodic SATs, "
            f"RK4 dt=0.2/{n - 1}")
    return Workload(sys_, 0.2 / (n - 1), n * n, desc, {"nx": n, "ny": n, "operator": "upwind9"}, cpu)


def build_c3(precision, cpu_steps):
    """24 Cartesian blocks (6 x 4 periodic layout over [-10, 10]^2, sizes
    drawn from seed 3 in [64, 1024]) coupled by Grid(j) interfaces, upwind4,
    vortex initial state, dt = 0.2·min over blocks of min(1/(nx-1), 1/(ny-1))."""
    from common import c3_dt, c3_layout
    from summationbyparts_amd import _lib, euler, multigrid, operators

    grids, nbs = c3_layout(seed=3)
    sys_ = multigrid.System(grids, [operators.Upwind4()] * len(grids))
    if precision == "mixed":
        sys_.set_precision(_lib.MIXED)
    for i, nb in enumerate(nbs):
        for d, j in nb.items():
            sys_.set_boundary(i, d, multigrid.Grid(j))
    vp = euler.VortexParameters([euler.Vortice(-1.0, 0.0, 0.5, 1.0)], 0.5)
    sys_.vortex(0.0, vp)
    dt = c3_dt(grids)
    pts = sum(g[0].size for g in grids)

    def cpu():
        # the oracle with one thread per block (the reference's distributed system)
        from oracle import oracle as O
        ref = O.System(grids, [(O.UPWIND4, O.UPWIND4)] * len(grids))
        for i, nb in enumerate(nbs):
            for d, j in nb.items():
                ref.set_bc(i, d, O.BC_GRID, j)
        ref.vortex(0.0, O.VortexParams.make([(-1.0, 0.0, 0.5, 1.0)], 0.5))
        threads = len(grids)
        t0 = time.perf_counter()
        ref.advance(dt, cpu_steps, threads)
        sec = time.perf_counter() - t0
        return _baseline(pts * cpu_steps, sec, min(threads, os.cpu_count() or 1),
                         f"{cpu_steps} RK4 steps of the 24-block layout, oracle, one thread per block "
                         f"({threads} threads on {os.cpu_count()} host cores)")

    desc = f"24 Cartesian blocks 6x4 over [-10,10]^2, sizes 64..1024 (seed 3), Grid interfaces, upwind4, vortex, dt={dt:.6g}"
    return Workload(sys_, dt, pts, desc, {"blocks": len(grids), "operator": "upwind4"}, cpu)


def run_adaptive(args):
    """advance_adaptive on the 4096² curvilinear upwind9 block (periodic,
    vortex + noise) with every step accepted (maxerr = inf): up to 40 embedded
    Bogacki-Shampine steps from a guess of 0.2/4095, the loop time growing by
    twice the step and the guess by 1.05 per step (after 40 steps the step is
    7x the first, a CFL number of about 0.4; many more would pass the
    stability limit)."""
    import torch
    from common import curvilinear, density_noise
    from summationbyparts_amd import euler, operators

    n = args.n
    X, Y = curvilinear(n, n)
    vp = euler.VortexParameters([euler.Vortice(-1.0, 0.0, 0.5, 1.0)], 0.5)
    s = euler.System(X, Y, operators.Upwind9())
    s.vortex(0.0, vp)
    q = s.field()
    q[0] += density_noise(1, 0, n, n)
    s.set_state(0, q)
    g0 = 0.2 / (n - 1)

    def span(k):
        # loop time reached after k accepted steps from guess g0, minus a
        # margin, so that exactly k steps are taken
        t, g = 0.0, g0
        for _ in range(k):
            t += 2.0 * g
            g *= 1.05
        return t - 0.5 * g0

    steps, warmup = min(args.steps, 40), min(args.warmup, 2)
    s.advance_adaptive(span(warmup), g0, np.inf)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = s.advance_adaptive(span(steps), g0, np.inf)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert abs(g - g0 * 1.05 ** steps) <= 1e-12 * g
    pts = n * n
    ms = wall * 1e3 / steps
    achieved = BYTES_PER_POINT_ADAPTIVE_STEP * pts / (ms * 1e-3) / 1e9
    out = {"metric": "Gpoint-embedded-RK-steps/s (advance_adaptive, Bogacki-Shampine)", "value": pts * steps / wall / 1e9,
           "unit": "Gpoint-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"advance_adaptive on the {n}x{n} curvilinear upwind9 block, periodic, all steps accepted",
                      "nx": n, "ny": n, "operator": "upwind9"},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None}}
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        m = 512
        Xs, Ys = curvilinear(m, m)
        ref = O.System([(Xs, Ys)], [(O.UPWIND9, O.UPWIND9)])
        ref.set_rhs_kind(O.RHS_UPWIND)
        ref.vortex(0.0, O.VortexParams.make([(-1.0, 0.0, 0.5, 1.0)], 0.5))
        t0 = time.perf_counter()
        ref.advance_adaptive(4.0 * 0.2 / (m - 1), 0.2 / (m - 1), np.inf)
        sec = time.perf_counter() - t0
        out["cpu_baseline"] = _baseline(m * m * 2, sec, 1, f"advance_adaptive, 2 accepted steps of a {m}x{m} block, oracle, 1 thread")
        out["cpu_baseline"]["unit"] = out["unit"]
    print(json.dumps(out), flush=True)


def two_to_one_layout(nc):
    """A coarse block [-5, 5] x [-5, 0] (nc x nc) under a fine block
    [-5, 5] x [0, 5] ((2nc-1) x (2nc-1)), Cartesian; the shared edges (and
    the periodic wrap in y) are Interpolate couplings, x is periodic (This)."""
    from common import cartesian
    nf = 2 * nc - 1
    xc, yc = cartesian(nc, nc)
    xf, yf = cartesian(nf, nf)
    return [(xc, (yc + 5.0) / 2.0 - 5.0), (xf, (yf + 5.0) / 2.0)]


def build_interp(precision, cpu_steps, nc=2048):
    """Two blocks with 2:1 interfaces: coarse nc x nc under fine (2nc-1)^2,
    upwind9, Interpolation9 on the shared and wrapped edges, vortex on the
    interface, dt = 0.2·(fine spacing)."""
    from summationbyparts_amd import _lib, euler, multigrid, operators

    grids = two_to_one_layout(nc)
    sys_ = multigrid.System(grids, [operators.Upwind9()] * 2)
    if precision == "mixed":
        sys_.set_precision(_lib.MIXED)
    I = operators.Interpolation9()
    for b, o in ((0, 1), (1, 0)):
        for d in (multigrid.NORTH, multigrid.SOUTH):
            sys_.set_boundary(b, d, multigrid.Interpolate(o, I))
    vp = euler.VortexParameters([euler.Vortice(-1.0, 0.0, 0.5, 1.0)], 0.5)
    sys_.vortex(0.0, vp)
    dt = 0.2 * 5.0 / (2 * nc - 2)
    pts = sum(g[0].size for g in grids)
    sample = 256

    def cpu():
        # the oracle on the same layout at nc = 256 (one thread per block)
        from oracle import oracle as O
        g = two_to_one_layout(sample)
        ref = O.System(g, [(O.UPWIND9, O.UPWIND9)] * 2)
        for b, o in ((0, 1), (1, 0)):
            for d in (O.NORTH, O.SOUTH):
                ref.set_bc_interpolate(b, d, o, O.INTERP9)
        ref.vortex(0.0, O.VortexParams.make([(-1.0, 0.0, 0.5, 1.0)], 0.5))
        n = sum(x.size for x, _ in g)
        steps = 4 * cpu_steps
        t0 = time.perf_counter()
        ref.advance(0.2 * 5.0 / (2 * sample - 2), steps, 2)
        sec = time.perf_counter() - t0
        return _baseline(n * steps, sec, 2,
                         f"{steps} RK4 steps of the same 2:1 layout at coarse {sample}^2 / fine {2 * sample - 1}^2, "
                         f"oracle, one thread per block")

    desc = (f"2:1 layout: coarse {nc}^2 under fine {2 * nc - 1}^2, Cartesian, upwind9, Interpolation9 interfaces "
            f"(periodic wrap), vortex, dt={dt:.6g}")
    return Workload(sys_, dt, pts, desc, {"blocks": 2, "operator": "upwind9", "interpolation": "Interpolation9"}, cpu)


def build_c5(precision, cpu_steps):
    """256 independent curvilinear vortex solves (seed 5: sizes in [48, 512]^2,
    Mach, eps, grid amplitude per solve, r* = 1, far-field vortex edges),
    upwind9, each with its own dt = 0.2·min(1/(nx-1), 1/(ny-1)) (dt scale per
    block, system step 1)."""
    from common import c5_batch, curvilinear
    from summationbyparts_amd import _lib, euler, operators

    batch = c5_batch()
    grids = [curvilinear(b["nx"], b["ny"], amp=b["amp"]) for b in batch]
    sys_ = euler.BlockSystem(grids, [operators.Upwind9()] * len(batch))
    if precision == "mixed":
        sys_.set_precision(_lib.MIXED)
    sys_.set_rhs_kind(_lib.RHS_UPWIND)
    params = []
    for i, b in enumerate(batch):
        vp = euler.VortexParameters([euler.Vortice(-1.0, 0.0, b["rstar"], b["eps"])], b["mach"])
        params.append(vp)
        for d in range(4):
            sys_.set_bc(i, d, _lib.BC_VORTEX, 0, vp)
        sys_.vortex(0.0, vp, block=i)
        sys_.set_dt_scale(i, b["dt"])
    pts = sum(b["nx"] * b["ny"] for b in batch)

    def cpu():
        # the oracle on the first 16 solves, each its own system, one thread
        from oracle import oracle as O
        npts, sec = 0, 0.0
        for b, (X, Y) in list(zip(batch, grids))[:16]:
            vo = O.VortexParams.make([(-1.0, 0.0, b["rstar"], b["eps"])], b["mach"])
            ref = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
            for d in range(4):
                ref.set_bc(0, d, O.BC_VORTEX, 0, vo)
            ref.vortex(0.0, vo)
            t0 = time.perf_counter()
            ref.advance(b["dt"], cpu_steps, 1)
            sec += time.perf_counter() - t0
            npts += b["nx"] * b["ny"] * cpu_steps
        return _baseline(npts, sec, 1, f"{cpu_steps} RK4 steps of the first 16 of the 256 solves, oracle, 1 thread")

    desc = "256 independent curvilinear vortex solves, sizes 48..512 (seed 5), own dt/Mach/eps each, far-field edges, upwind9"
    return Workload(sys_, 1.0, pts, desc, {"blocks": len(batch), "operator": "upwind9"}, cpu)


def mixed_leg(n, steps, warmup, fp64_value):
    """The metric's second half on the default workload: the mixed form's
    throughput over the same timed steps, and its accuracy against the FP64
    form after 1000 RK4 steps of the same block, vortex and noise with the
    analytic vortex as far-field data on all edges (the mixed-precision
    configuration; the grid's periodic seam is non-conforming and a periodic
    run turns unstable after about 800 steps in either precision)."""
    import torch
    from common import curvilinear, density_noise
    from summationbyparts_amd import _lib, euler, operators

    w = build_c2(n, "mixed", 0)
    s = w.sys
    s.advance_steps(w.dt, warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms_timed = s.advance_timed(w.dt, steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = n * n * steps / wall / 1e9
    ms_stage = ms_timed / (4 * steps)
    s.close()
    X, Y = curvilinear(n, n)
    vp = euler.VortexParameters([euler.Vortice(-1.0, 0.0, 0.5, 1.0)], 0.5)
    out, drift = {}, {}
    for prec in ("fp64", "mixed"):
        c = euler.System(X, Y, operators.Upwind9())
        c.set_precision(_lib.MIXED if prec == "mixed" else _lib.FP64)
        for d in range(4):
            c.set_bc(0, d, _lib.BC_VORTEX, 0, vp)
        c.vortex(0.0, vp)
        q = c.field()
        q[0] += density_noise(1, 0, n, n)
        c.set_state(0, q)
        e0 = c.energy()
        c.advance_n(0.2 / (n - 1), 1000)
        drift[prec] = (c.energy() - e0) / e0
        out[prec] = c.field()
        c.close()
    a, b = out["mixed"], out["fp64"]
    rel_l2 = float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
    per_comp = [float(np.linalg.norm(a[k] - b[k]) / np.linalg.norm(b[k])) for k in range(4)]
    return {"value": value, "unit": UNIT, "ratio_vs_fp64": value / fp64_value,
            "stage_launch_ms": ms_stage,
            "roofline_frac": BYTES_PER_POINT_STEP["mixed"] / 4.0 * n * n / (ms_stage * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "rel_l2_vs_fp64": rel_l2, "rel_l2_per_component": per_comp, "tolerance": 1e-8,
            "accuracy_run": "1000 RK4 steps, far-field vortex edges, same initial state",
            "energy_drift": drift}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--workload", choices=["c2", "c3", "c5", "interp", "adaptive"], default="c2",
                    help="c2: one 4096^2 curvilinear block (default); c3: 24-block Grid layout; c5: 256-solve batch; "
                         "interp: 2048^2 block under a 4095^2 block, 2:1 Interpolate interfaces; "
                         "adaptive: advance_adaptive (embedded Bogacki-Shampine) on the 4096^2 block")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="default workload only: skip the mixed-form leg (throughput and 1000-step error vs FP64)")
    ap.add_argument("--tile", type=int, default=-1, help="force a tile shape index (-1: autotune)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 2 GiB device copy (reference kernel for HBM byte counters)")
    ap.add_argument("--counters", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="summarize rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this bench (with --tile) "
                         "into HBM bytes per stage launch, calibrated by two runs of --calibrate-copy")
    args = ap.parse_args()
    if args.counters:
        print(json.dumps(summarize_counters(*args.counters, args.n, args.tile, args.precision)))
        return
    if args.calibrate_copy:
        import torch

        a = torch.ones(1 << 28, dtype=torch.float64, device="cuda")  # 2 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

        def rate(fn, nbytes, reps=10):
            fn()
            torch.cuda.synchronize()
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            return nbytes / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9

        n = a.numel() * 8
        q = a.numel() // 4
        c, d, e, o = a[:q], a[q:2 * q], a[2 * q:3 * q], b[:q]
        out = {"copy_bytes_read": n, "copy_bytes_written": n,
               "copy_gbs": rate(lambda: b.copy_(a), 2 * n),
               "sum_gbs": rate(lambda: a.sum(), n),
               "addcmul_3r1w_gbs": rate(lambda: torch.addcmul(c, d, e, out=o), 4 * q * 8)}
        print(json.dumps(out))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    from summationbyparts_amd import euler

    euler.set_gamma(1.4)
    if args.workload == "adaptive":
        run_adaptive(args)
        return
    if args.workload == "c2":
        w = build_c2(args.n, args.precision, args.cpu_steps)
    elif args.workload == "c3":
        w = build_c3(args.precision, args.cpu_steps)
    elif args.workload == "interp":
        w = build_interp(args.precision, args.cpu_steps)
    else:
        w = build_c5(args.precision, args.cpu_steps)
    sys_ = w.sys
    if args.tile >= 0:
        sys_.set_tile_shape(args.tile)
    dt = w.dt
    sys_.advance_steps(dt, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # HIP events on the library's stream bracket the K steps (graph replays)
    ms_timed = sys_.advance_timed(dt, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    upload_ms = None
    if args.workload == "c2":
        # host -> device upload of one state through the boundary (not part of
        # the timed region: a run pays it once)
        q = sys_.field()
        t1 = time.perf_counter()
        sys_.set_state(0, q)
        torch.cuda.synchronize()
        upload_ms = (time.perf_counter() - t1) * 1e3
    # mean duration of one stage launch (the stage kernel over the interior
    # tiles, concurrent with its edge-tile instantiation) over the timed
    # region: 4 launches per step plus a 1-thread time update
    ms_stage = ms_timed / (4 * args.steps)
    tile, tile_index = sys_.tile_shape(), sys_.tile_index()
    pts = w.points
    value = pts * args.steps / wall / 1e9
    bytes_per_launch = BYTES_PER_POINT_STEP[args.precision] / 4.0 * pts
    achieved = bytes_per_launch / (ms_stage * 1e-3) / 1e9
    # HBM bytes per launch measured with rocprofv3 counters for this
    # workload and tile shape, if recorded
    traffic = None
    tfile = TRAFFIC_FILES[args.precision]
    if os.path.exists(tfile) and args.workload == "c2":
        with open(tfile) as f:
            tr = json.load(f)
        # the mixed form runs shapes 0 and 1 as the same kernel (it always
        # keeps flux and α·q together in LDS)
        ti = 1 if (args.precision == "mixed" and tile_index == 0) else tile_index
        if tr.get("n") == args.n and tr.get("precision", "fp64") == args.precision:
            for e in tr.get("tiles", []):
                if e.get("tile_index") == ti:
                    traffic = e.get("hbm_bytes_per_launch")
    cfg = {"workload": w.desc, **w.cfg, "precision": args.precision, "tile": list(tile),
           "tile_index": tile_index, "event_ms_per_step": ms_timed / args.steps, "stage_launch_ms": ms_stage}
    if upload_ms is not None:
        cfg["state_upload_ms"] = upload_ms
    out = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-rhs/f64-state",
        "data": "synthetic",
        "config": cfg,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "reference_schedule": {"bytes_per_point_step": REFERENCE_SCHEDULE_BYTES,
                               "equivalent_gbs": REFERENCE_SCHEDULE_BYTES / 4.0 * pts / (ms_stage * 1e-3) / 1e9,
                               "frac": REFERENCE_SCHEDULE_BYTES / 4.0 * pts / (ms_stage * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    if args.workload == "c2" and args.precision == "fp64" and args.tile < 0 and not args.no_mixed:
        w.sys.close()
        out["mixed"] = mixed_leg(args.n, args.steps, args.warmup, value)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = w.cpu()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
