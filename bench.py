"""Benchmark: Euler RK4 step throughput on one MI355X.

Headline workload: one curvilinear 4096x4096 block, upwind9 with upwind
dissipation, isentropic vortex (centre (-1, 0), r* 0.5, eps 1, Mach 0.5) plus
seeded +-1e-3 density noise, periodic SATs, dt = 0.2/4095, FP64. A step is one
classical RK4 step (four stage launches) over the whole block with the state
resident in HBM.

The default run prints ONE JSON line: the headline throughput in
Gpoint-RK4-steps/s with its HBM roofline (algorithmic bytes per stage launch
over the mean launch duration from HIP events on the library's stream), the
mixed form on the same block (throughput and its error against FP64 after
1000 steps), the CPU oracle timed on this host (median of 3, one thread, plus
an all-core variant), and under "workloads" the same figures for the other
configurations:
  c1        101x101 Cartesian, SBP8, far-field vortex edges, 100 steps
  c3        24 Cartesian blocks (sizes 64..1024), Grid interfaces, upwind4, 500 steps
  c5_fp64   256 independent curvilinear vortex solves (48..512), 300 steps
  c5_mixed  the same batch in the mixed form, with its error against c5_fp64
  interp    a 2048^2 block under a 4095^2 block, 2:1 Interpolate interfaces
  adaptive  advance_adaptive (embedded Bogacki-Shampine) on the 4096^2 block
`--workload NAME [--precision mixed]` prints the line of one workload alone;
`--main-only` skips the "workloads" object.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "Gpoint-RK4-steps/s + % HBM roofline; mixed vs FP64 at ≤1e-8 rel L2 err"
UNIT = "Gpoint-RK4-steps/s"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per point per RK4 step as the headline metric defines
# them: the state, the 5 FP64 metric planes and the y_n / accumulator /
# stage-input traffic of a fused schedule of the reference's step (672 B);
# the mixed form with FP32 stage inputs (576 B).
BYTES_PER_POINT_STEP = {"fp64": 672.0, "mixed": 576.0}
# This kernel's own schedule moves fewer bytes. FP64 (J formed from the four
# metric products, so 4 metric planes): stage 1 reads y_n, metrics, writes
# Y2 (96 B); stage 2 reads Y2, metrics, y_n, writes Y3 and W = Y2 + 2 Y3 - y_n
# (160 B); stage 3 reads Y3, metrics, y_n, writes Y4 (128 B); stage 4 reads
# Y4, metrics, W and writes y_{n+1} (128 B): 512 B.
# Mixed: the stage inputs are y_n + Δ_s with FP32 increments Δ_s, and the
# last stage sums y_n and the increments in FP64: stage 1 reads y_n, metrics,
# writes Δ_1 (80 B); stages 2-3 read y_n, Δ_s, metrics and write Δ_{s+1}
# (96 B each); stage 4 reads y_n, Δ_3, metrics, Δ_1, Δ_2 and writes y_{n+1}
# (144 B): 416 B.
OWN_BYTES_PER_POINT_STEP = {"fp64": 512.0, "mixed": 416.0}
# advance_adaptive's embedded Bogacki-Shampine trial step, fused into four
# stage launches: stage 1 reads y_n and the metric pairs and writes Y2
# (96 B); stage 2 reads Y2, the metrics and y_n and writes Y3 (128 B);
# stage 3 reads Y3, the metrics, y_n and Y2 and writes the solution (160 B);
# stage 4 reads the solution, the metrics, y_n and Y2 and writes only the
# per-tile error sums (128 B): 512 B per point per trial step (the accepted
# solution is copied back once per call, not per step).
BYTES_PER_POINT_ADAPTIVE_STEP = 512.0
MIXED_TOL = 1e-8
# HBM bytes per stage launch from rocprofv3 counters (bench.py --counters),
# per workload and precision
TRAFFIC_FILES = {("c2", "fp64"): "stage_kernel_upwind9_4096_traffic.json",
                 ("c2", "mixed"): "stage_kernel_upwind9_4096_mixed_traffic.json",
                 ("c3", "fp64"): "stage_kernel_c3_24_blocks_upwind4_traffic.json",
                 ("c5", "fp64"): "stage_kernel_c5_256_solves_upwind9_traffic.json",
                 ("c5", "mixed"): "stage_kernel_c5_256_solves_upwind9_mixed_traffic.json"}


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


def _step_kernel(name):
    """Dispatches that belong to a stage launch: the stage kernel (interior
    and edge-tile instantiations), the per-stage edge SATs and Eval edges."""
    return any(k in name for k in ("stage_kernel", "edge_sat_kernel", "eval_edges_kernel", "mix_ref_kernel"))


def _is_interior_dispatch(name):
    """A stage-kernel dispatch of the interior instantiation: the EDGE template
    argument (the tenth: OX, OY, UP, TX, TY, RX, RY, VAR, MIX, EDGE, ADAPT)
    is false."""
    if "stage_kernel<" not in name:
        return False
    args = name.split("(")[0].split("<", 1)[1].rsplit(">", 1)[0].split(", ")
    return len(args) >= 10 and args[9] == "false"


def summarize_counters(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, workload, tile_index, precision, points):
    """HBM bytes per stage launch from FETCH_SIZE / WRITE_SIZE (KiB per
    dispatch) of a run of this bench, scaled by what the same counters read
    for the 2 GiB device copy of `--calibrate-copy`. A stage launch is one
    interior-tile dispatch of the stage kernel with the edge-tile and
    edge-SAT dispatches of the same stage; the first step (cold caches, the
    first 4 launches) is left out and the rest averaged per launch and per
    RK4 stage."""
    copy_bytes = float(1 << 31)

    def factor(path, counter):
        vals = sorted(v for _, k, v in _counter_rows(path, counter) if "copyBuffer" in k)
        return copy_bytes / (vals[len(vals) // 2] * 1024.0)

    ff, wf = factor(copy_fetch_csv, "FETCH_SIZE"), factor(copy_write_csv, "WRITE_SIZE")

    def per_launch(path, counter, scale):
        rows = [(k, v * 1024.0 * scale) for _, k, v in _counter_rows(path, counter) if _step_kernel(k)]
        interior = [i for i, (k, _) in enumerate(rows) if _is_interior_dispatch(k)]
        if len(interior) < 8 or len(interior) % 4:
            raise SystemExit(f"{path}: {len(interior)} interior stage dispatches (want whole steps, at least 2)")
        kernel = rows[interior[0]][0]
        # launch j: from the dispatch after launch j-1's interior dispatch
        # up to and including its own (the aux stream's SAT and edge tiles are
        # submitted before the interior tiles of the same stage)
        bounds = [-1] + interior
        launches = [sum(v for _, v in rows[bounds[j] + 1:bounds[j + 1] + 1]) for j in range(len(interior))]
        stages = [[] for _ in range(4)]
        for i, v in enumerate(launches[4:]):
            stages[i % 4].append(v)
        return kernel, [sum(s) / len(s) for s in stages]

    kernel, fetch = per_launch(fetch_csv, "FETCH_SIZE", ff)
    _, write = per_launch(write_csv, "WRITE_SIZE", wf)
    return {"workload": workload, "tile_index": tile_index, "tile": list(_tile_of(tile_index)),
            "kernel": kernel.split("(")[0], "points": points,
            "hbm_bytes_per_launch": (sum(fetch) + sum(write)) / 4.0,
            "algorithmic_bytes_per_launch": BYTES_PER_POINT_STEP[precision] / 4.0 * points,
            "own_schedule_bytes_per_launch": OWN_BYTES_PER_POINT_STEP[precision] / 4.0 * points,
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


def host_threads():
    """Host threads for the all-core CPU variant: the cores this process may
    run on, at most the 16 a one-GPU share holds (OMP_NUM_THREADS caps it
    where set)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(n, 16, cap))


class Workload:
    """A system ready to step, its step size, point count, description and
    CPU-oracle samples: `cpu(inner)` times one bounded run of the oracle
    (inner = threads inside each block; 1 is the reference's schedule) and
    returns (point-steps, seconds); `cpu_sample` says what the run is;
    `all_core` says whether an all-core variant differs from the reference's
    schedule."""

    def __init__(self, sys_, dt, points, desc, cfg, cpu, cpu_sample, cpu_cores, all_core=True):
        self.sys, self.dt, self.points, self.desc, self.cfg = sys_, dt, points, desc, cfg
        self.cpu, self.cpu_sample, self.cpu_cores, self.all_core = cpu, cpu_sample, cpu_cores, all_core


def cpu_baseline(w, unit=UNIT, runs=3):
    """The oracle on this host: the median of `runs` timed runs with the
    reference's schedule (one thread per block), and, where it differs, the
    median with every host core working inside each block (labelled
    "stronger than reference"; results are bit-identical to one thread)."""
    from oracle import oracle as O

    def median_of(inner):
        O.set_inner_threads(inner)
        try:
            vals = []
            for _ in range(runs):
                ps, sec = w.cpu(inner)
                vals.append(ps / sec / 1e9)
        finally:
            O.set_inner_threads(1)
        return statistics.median(vals), vals

    v, vals = median_of(1)
    out = {"value": v, "unit": unit, "cores": w.cpu_cores, "kind": "port",
           "sample": w.cpu_sample + f"; median of {runs} runs", "runs": vals,
           "host_cores_visible": os.cpu_count()}
    if w.all_core:
        T = host_threads()
        va, valsa = median_of(T)
        out["all_core"] = {"value": va, "unit": unit, "cores": T * w.cpu_cores, "kind": "port",
                           "label": "stronger than reference",
                           "sample": f"the same runs with {T} threads inside each block; median of {runs} runs",
                           "runs": valsa}
    return out


def stored_traffic(workload, precision, tile_index, points):
    """HBM bytes per stage launch recorded by a rocprofv3 --pmc run of this
    bench for the same workload, precision and tile shape, with its source;
    (None, reason) where none is recorded."""
    name = TRAFFIC_FILES.get((workload, precision))
    if name is None:
        return None, "not measured for this workload"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, f"not measured (profiles/{name} absent)"
    with open(path) as f:
        tr = json.load(f)
    # the mixed form runs shapes 0 and 1 as the same kernel (it always keeps
    # flux and α·q together in LDS)
    ti = 1 if (precision == "mixed" and tile_index == 0) else tile_index
    for e in tr.get("tiles", []):
        if e.get("tile_index") == ti and e.get("points", points) == points:
            return e["hbm_bytes_per_launch"], (f"stored: profiles/{name}, rocprofv3 --pmc FETCH_SIZE then WRITE_SIZE "
                                              f"of this bench, tile index {ti}, copy-calibrated")
    return None, f"not measured for tile index {ti} (profiles/{name} holds other tiles)"


def roofline(precision, points, ms_stage, traffic=None, traffic_source="not measured for this workload",
             bytes_per_point_step=None, definition=None):
    """HBM roofline of one stage launch: the headline metric's algorithmic
    bytes (672 B FP64 / 576 B mixed per point-step, a quarter per launch)
    over the mean launch duration, and this kernel's own schedule beside."""
    bpp = bytes_per_point_step or BYTES_PER_POINT_STEP[precision]
    per_launch = bpp / 4.0 * points
    achieved = per_launch / (ms_stage * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
           "traffic": traffic, "traffic_source": traffic_source, "bytes_per_point_step": bpp,
           "algorithmic_bytes_per_launch": per_launch,
           "bytes_definition": definition or (f"{bpp:.0f} B per point per RK4 step (the metric's definition: state, "
                                              f"5 metric planes, y_n/accumulator/stage inputs), a quarter per launch")}
    if bytes_per_point_step is None:
        own = OWN_BYTES_PER_POINT_STEP[precision] / 4.0 * points / (ms_stage * 1e-3) / 1e9
        out["own_schedule"] = {"bytes_per_point_step": OWN_BYTES_PER_POINT_STEP[precision], "achieved": own,
                               "frac": own / HBM_PEAK_GBS}
    if traffic is not None:
        out["traffic_gbs"] = traffic / (ms_stage * 1e-3) / 1e9
    return out


def time_steps(w, steps, warmup, tile=-1, schedule=-1):
    """Warm up, then time `steps` graph-replayed steps between a device
    synchronization on both sides (HIP events on the library stream give the
    per-launch duration)."""
    import torch
    s = w.sys
    if tile >= 0:
        s.set_tile_shape(tile)
    if schedule >= 0:
        s.set_schedule(schedule)
    s.advance_steps(w.dt, warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms_timed = s.advance_timed(w.dt, steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # mean duration of one stage launch (edge SATs and edge tiles on the
    # second stream, concurrent with the interior tiles) over the timed
    # region: 4 launches per step (the first two also carry the time update)
    return {"value": w.points * steps / wall / 1e9, "wall": wall, "ms_per_step": wall * 1e3 / steps,
            "event_ms_per_step": ms_timed / steps, "stage_launch_ms": ms_timed / (4 * steps),
            "tile": list(s.tile_shape()), "tile_index": s.tile_index(),
            "schedule": ["split", "merged", "merged after the edge-SAT kernel"][s.schedule()],
            "tile_order": ["XCD bands", "row-major"][s.tile_order()]}


def line(name, w, precision, t, steps, warmup, traffic_key=None, metric=METRIC):
    """A JSON object of the bench contract for one timed workload."""
    traffic, source = (None, "not measured for this workload")
    if traffic_key is not None:
        traffic, source = stored_traffic(traffic_key, precision, t["tile_index"], w.points)
    cfg = {"workload": w.desc, **w.cfg, "precision": precision, "tile": t["tile"], "tile_index": t["tile_index"],
           "schedule": t["schedule"], "tile_order": t["tile_order"],
           "event_ms_per_step": t["event_ms_per_step"], "stage_launch_ms": t["stage_launch_ms"],
           "points": w.points}
    return {"metric": metric, "value": t["value"], "unit": UNIT, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": t["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64" if precision == "fp64" else "f32-rhs/f64-state", "data": "synthetic", "config": cfg,
            "roofline": roofline(precision, w.points, t["stage_launch_ms"], traffic, source)}


# ---------------------------------------------------------------------------
# workloads


def build_c1(precision):
    """101x101 Cartesian on [-5, 5]^2, SBP8 (traditional), the vortex
    (centre (-1, 0), r* 0.5, eps 1, Mach 0.5) as far-field data on all four
    edges, dt = 0.2/100."""
    from common import cartesian
    from summationbyparts_amd import _lib, euler, multigrid, operators

    n = 101
    X, Y = cartesian(n, n)
    vp = euler.VortexParameters([euler.Vortice(-1.0, 0.0, 0.5, 1.0)], 0.5)
    sys_ = multigrid.System([(X, Y)], [operators.SBP8()])
    if precision == "mixed":
        sys_.set_precision(_lib.MIXED)
    for d in range(4):
        sys_.set_boundary(0, d, multigrid.Vortex(vp))
    sys_.vortex(0.0, vp)
    dt = 0.2 / (n - 1)

    def cpu(inner):
        from oracle import oracle as O
        vo = O.VortexParams.make([(-1.0, 0.0, 0.5, 1.0)], 0.5)
        ref = O.System([(X, Y)], [(O.SBP8, O.SBP8)])
        for d in range(4):
            ref.set_bc(0, d, O.BC_VORTEX, 0, vo)
        ref.vortex(0.0, vo)
        t0 = time.perf_counter()
        ref.advance(dt, 100, 1)
        return n * n * 100, time.perf_counter() - t0

    return Workload(sys_, dt, n * n, f"single Cartesian block {n}x{n}, SBP8, far-field vortex SATs, RK4 dt=0.002",
                    {"nx": n, "ny": n, "operator": "sbp8"}, cpu,
                    "the 100 RK4 steps of the 101x101 SBP8 block, oracle (C++ restatement), 1 thread", 1,
                    all_core=False)


def c2_state(n):
    from common import curvilinear, density_noise
    from summationbyparts_amd import euler, operators
    X, Y = curvilinear(n, n)
    s = euler.System(X, Y, operators.Upwind9())
    s.vortex(0.0, euler.VortexParameters([euler.Vortice(-1.0, 0.0, 0.5, 1.0)], 0.5))
    q = s.field()
    q[0] += density_noise(1, 0, n, n)
    return X, Y, s, q


def build_c2(n, precision, cpu_steps=2):
    """One curvilinear n x n block, upwind9 + dissipation, vortex + 1e-3
    density noise, periodic SATs, dt = 0.2/(n-1)."""
    from summationbyparts_amd import _lib

    X, Y, sys_, q = c2_state(n)
    if precision == "mixed":
        sys_.set_precision(_lib.MIXED)
    sys_.set_state(0, q)
    sys_.set_rhs_kind(_lib.RHS_UPWIND)
    ref = {}

    def cpu(inner):
        # one block: the reference's one thread per block is one thread
        from oracle import oracle as O
        if "s" not in ref:
            ref["s"] = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
            ref["s"].set_rhs_kind(O.RHS_UPWIND)
            ref["s"].set_state(0, q)
        t0 = time.perf_counter()
        ref["s"].advance(0.2 / (n - 1), cpu_steps, 1)
        return n * n * cpu_steps, time.perf_counter() - t0

    desc = (f"single curvilinear block {n}x{n}, upwind9 + dissipation, vortex + 1e-3 density noise, periodic SATs, "
            f"RK4 dt=0.2/{n - 1}")
    return Workload(sys_, 0.2 / (n - 1), n * n, desc, {"nx": n, "ny": n, "operator": "upwind9"}, cpu,
                    f"{cpu_steps} RK4 steps of the {n}x{n} upwind9 curvilinear block, oracle (C++ restatement), "
                    f"1 thread", 1)


def build_c3(precision, cpu_steps=3):
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
    ref = {}

    def cpu(inner):
        # one thread per block (the reference's distributed system)
        from oracle import oracle as O
        if "s" not in ref:
            r = O.System(grids, [(O.UPWIND4, O.UPWIND4)] * len(grids))
            for i, nb in enumerate(nbs):
                for d, j in nb.items():
                    r.set_bc(i, d, O.BC_GRID, j)
            r.vortex(0.0, O.VortexParams.make([(-1.0, 0.0, 0.5, 1.0)], 0.5))
            ref["s"] = r
        t0 = time.perf_counter()
        ref["s"].advance(dt, cpu_steps, len(grids))
        return pts * cpu_steps, time.perf_counter() - t0

    desc = f"24 Cartesian blocks 6x4 over [-10,10]^2, sizes 64..1024 (seed 3), Grid interfaces, upwind4, vortex, dt={dt:.6g}"
    ncpu = os.cpu_count() or 1
    return Workload(sys_, dt, pts, desc, {"blocks": len(grids), "operator": "upwind4"}, cpu,
                    f"{cpu_steps} RK4 steps of the 24-block layout, oracle, one thread per block "
                    f"(24 threads; the reference's schedule already puts every block on its own core)",
                    min(len(grids), ncpu), all_core=False)


def two_to_one_layout(nc):
    """A coarse block [-5, 5] x [-5, 0] (nc x nc) under a fine block
    [-5, 5] x [0, 5] ((2nc-1) x (2nc-1)), Cartesian; the shared edges (and
    the periodic wrap in y) are Interpolate couplings, x is periodic (This)."""
    from common import cartesian
    nf = 2 * nc - 1
    xc, yc = cartesian(nc, nc)
    xf, yf = cartesian(nf, nf)
    return [(xc, (yc + 5.0) / 2.0 - 5.0), (xf, (yf + 5.0) / 2.0)]


def build_interp(precision, nc=2048, sample=256, cpu_steps=12):
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
    ref = {}

    def cpu(inner):
        # the oracle on the same layout at coarse 256^2 (one thread per block)
        from oracle import oracle as O
        if "s" not in ref:
            g = two_to_one_layout(sample)
            r = O.System(g, [(O.UPWIND9, O.UPWIND9)] * 2)
            for b, o in ((0, 1), (1, 0)):
                for d in (O.NORTH, O.SOUTH):
                    r.set_bc_interpolate(b, d, o, O.INTERP9)
            r.vortex(0.0, O.VortexParams.make([(-1.0, 0.0, 0.5, 1.0)], 0.5))
            ref["s"], ref["n"] = r, sum(x.size for x, _ in g)
        t0 = time.perf_counter()
        ref["s"].advance(0.2 * 5.0 / (2 * sample - 2), cpu_steps, 2)
        return ref["n"] * cpu_steps, time.perf_counter() - t0

    desc = (f"2:1 layout: coarse {nc}^2 under fine {2 * nc - 1}^2, Cartesian, upwind9, Interpolation9 interfaces "
            f"(periodic wrap), vortex, dt={dt:.6g}")
    return Workload(sys_, dt, pts, desc, {"blocks": 2, "operator": "upwind9", "interpolation": "Interpolation9"}, cpu,
                    f"{cpu_steps} RK4 steps of the same 2:1 layout at coarse {sample}^2 / fine {2 * sample - 1}^2, "
                    f"oracle, one thread per block", 2)


def build_c5(precision, sample=32, cpu_steps=3):
    """256 independent curvilinear vortex solves (seed 5: sizes in [48, 512]^2,
    Mach, eps, grid amplitude per solve, r* = 1, far-field vortex edges),
    upwind9, each with its own dt = 0.2·min(1/(nx-1), 1/(ny-1)) (dt scale per
    block, system step 1)."""
    from concurrent.futures import ThreadPoolExecutor

    from common import c5_batch, curvilinear
    from summationbyparts_amd import _lib, euler, operators

    batch = c5_batch()
    grids = [curvilinear(b["nx"], b["ny"], amp=b["amp"]) for b in batch]
    sys_ = euler.BlockSystem(grids, [operators.Upwind9()] * len(batch))
    if precision == "mixed":
        sys_.set_precision(_lib.MIXED)
    sys_.set_rhs_kind(_lib.RHS_UPWIND)
    for i, b in enumerate(batch):
        vp = euler.VortexParameters([euler.Vortice(-1.0, 0.0, b["rstar"], b["eps"])], b["mach"])
        for d in range(4):
            sys_.set_bc(i, d, _lib.BC_VORTEX, 0, vp)
        sys_.vortex(0.0, vp, block=i)
        sys_.set_dt_scale(i, b["dt"])
    pts = sum(b["nx"] * b["ny"] for b in batch)
    T = host_threads()
    ref = {}

    def cpu(inner):
        # the first `sample` solves, each its own oracle system, on a pool of
        # T threads (one solve per task, as a rayon pool over the batch)
        from oracle import oracle as O
        if "s" not in ref:
            systems = []
            for b, (X, Y) in list(zip(batch, grids))[:sample]:
                vo = O.VortexParams.make([(-1.0, 0.0, b["rstar"], b["eps"])], b["mach"])
                r = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
                r.set_rhs_kind(O.RHS_UPWIND)
                for d in range(4):
                    r.set_bc(0, d, O.BC_VORTEX, 0, vo)
                r.vortex(0.0, vo)
                systems.append((r, b["dt"], b["nx"] * b["ny"]))
            ref["s"] = systems
        t0 = time.perf_counter()
        with ThreadPoolExecutor(T) as ex:
            list(ex.map(lambda e: e[0].advance(e[1], cpu_steps, 1), ref["s"]))
        return sum(e[2] for e in ref["s"]) * cpu_steps, time.perf_counter() - t0

    desc = "256 independent curvilinear vortex solves, sizes 48..512 (seed 5), own dt/Mach/eps each, far-field edges, upwind9"
    return Workload(sys_, 1.0, pts, desc, {"blocks": len(batch), "operator": "upwind9"}, cpu,
                    f"{cpu_steps} RK4 steps of the first {sample} of the 256 solves, oracle, a pool of {T} threads "
                    f"(one solve per task: the reference's schedule already uses every core)", T, all_core=False), batch


def rel_errors(a, b):
    """Relative L2 and max-norm errors of a against b, whole state and per
    component (rho, rho u, rho v, e); per component both over the
    component's own norm and over the whole state's norm (the component's
    share of the whole-state error: rho v vanishes outside the vortex, so
    its own norm is small)."""
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    return {"rel_l2": float(np.linalg.norm((a - b).ravel()) / nb),
            "l2_per_component_over_state": [float(np.linalg.norm(a[k] - b[k]) / nb) for k in range(4)],
            "rel_max": float(np.abs(a - b).max() / np.abs(b).max()),
            "rel_l2_per_component": [float(np.linalg.norm(a[k] - b[k]) / np.linalg.norm(b[k])) for k in range(4)],
            "rel_max_per_component": [float(np.abs(a[k] - b[k]).max() / np.abs(b[k]).max()) for k in range(4)]}


def mixed_leg(n, steps, warmup, fp64_value):
    """The metric's second half on the headline block: the mixed form's
    throughput over the same timed steps, and its accuracy against the FP64
    form after 1000 RK4 steps of the same block, vortex and noise with the
    analytic vortex as far-field data on all edges (the grid's periodic seam
    is non-conforming and a periodic run turns unstable after about 800
    steps in either precision)."""
    from summationbyparts_amd import _lib, euler

    w = build_c2(n, "mixed")
    t = time_steps(w, steps, warmup)
    w.sys.close()
    X, Y, base, q = c2_state(n)
    base.close()
    vp = euler.VortexParameters([euler.Vortice(-1.0, 0.0, 0.5, 1.0)], 0.5)
    out, drift = {}, {}
    for prec in ("fp64", "mixed"):
        from summationbyparts_amd import operators
        c = euler.System(X, Y, operators.Upwind9())
        c.set_precision(_lib.MIXED if prec == "mixed" else _lib.FP64)
        for d in range(4):
            c.set_bc(0, d, _lib.BC_VORTEX, 0, vp)
        c.set_state(0, q)
        e0 = c.energy()
        c.advance_n(0.2 / (n - 1), 1000)
        drift[prec] = (c.energy() - e0) / e0
        out[prec] = c.field()
        c.close()
    err = rel_errors(out["mixed"], out["fp64"])
    ratio = t["value"] / fp64_value
    return {"value": t["value"], "unit": UNIT, "ratio_vs_fp64": ratio, "steps": steps,
            "ms_per_step": t["ms_per_step"], "stage_launch_ms": t["stage_launch_ms"], "tile": t["tile"],
            "tile_index": t["tile_index"],
            "roofline": roofline("mixed", n * n, t["stage_launch_ms"], *stored_traffic("c2", "mixed", t["tile_index"],
                                                                                        n * n)),
            **err, "tolerance": MIXED_TOL, "within_tolerance": err["rel_l2"] <= MIXED_TOL,
            "components_within_tolerance": [e <= MIXED_TOL for e in err["l2_per_component_over_state"]],
            "components_definition": "||a_c - b_c|| / ||b|| (each component's share of the whole-state error)",
            "byte_ratio_ceiling": OWN_BYTES_PER_POINT_STEP["fp64"] / OWN_BYTES_PER_POINT_STEP["mixed"],
            "accuracy_run": "1000 RK4 steps, far-field vortex edges, same initial state, against the FP64 form",
            "energy_drift": drift}


def run_c5_pair(steps, warmup, no_cpu):
    """C5 in FP64 and in the mixed form (the same batch, steps and timing),
    and the mixed form's error against FP64 per solve after the timed steps."""
    import torch
    from summationbyparts_amd import euler, operators

    res, states, verr = {}, {}, {}
    base = None
    for prec in ("fp64", "mixed"):
        w, batch = build_c5(prec)
        t = time_steps(w, steps, warmup)
        res[prec] = line("c5", w, prec, t, steps, warmup, traffic_key="c5")
        states[prec] = [w.sys.get_state(i) for i in range(len(batch))]
        # System::error of every solve against its own analytic vortex
        verr[prec] = [w.sys.error_vortex(euler.VortexParameters([euler.Vortice(-1.0, 0.0, b["rstar"], b["eps"])],
                                                                 b["mach"]), block=i) for i, b in enumerate(batch)]
        if prec == "fp64" and not no_cpu:
            base = cpu_baseline(w)
        w.sys.close()
    # the mixed-vs-FP64 difference in the H-norm of the error, against the
    # discretisation error of the FP64 solve
    op = operators.Upwind9()
    ratio_disc, change = 0.0, 0.0
    for i in range(len(states["fp64"])):
        a = torch.from_numpy(states["mixed"][i]).cuda()
        b = torch.from_numpy(states["fp64"][i]).cuda()
        ratio_disc = max(ratio_disc, euler.h2_err(a, b, op) / verr["fp64"][i])
        change = max(change, abs(verr["mixed"][i] - verr["fp64"][i]) / verr["fp64"][i])
    worst = {"rel_l2": 0.0, "rel_max": 0.0}
    comp_l2, comp_max = [0.0] * 4, [0.0] * 4
    for a, b in zip(states["mixed"], states["fp64"]):
        e = rel_errors(a, b)
        worst["rel_l2"] = max(worst["rel_l2"], e["rel_l2"])
        worst["rel_max"] = max(worst["rel_max"], e["rel_max"])
        comp_l2 = [max(x, y) for x, y in zip(comp_l2, e["rel_l2_per_component"])]
        comp_max = [max(x, y) for x, y in zip(comp_max, e["rel_max_per_component"])]
    res["mixed"]["vs_fp64"] = {"ratio": res["mixed"]["value"] / res["fp64"]["value"],
                               "worst_solve_rel_l2": worst["rel_l2"], "worst_solve_rel_max": worst["rel_max"],
                               "worst_solve_rel_l2_per_component": comp_l2,
                               "worst_solve_rel_max_per_component": comp_max, "tolerance": MIXED_TOL,
                               "within_tolerance": worst["rel_l2"] <= MIXED_TOL,
                               "worst_solve_diff_over_discretisation_error": ratio_disc,
                               "worst_solve_rel_change_of_vortex_error": change,
                               "accuracy_run": f"{warmup + steps} RK4 steps of every solve, same initial states"}
    if base is not None:
        res["fp64"]["cpu_baseline"] = base
        res["mixed"]["cpu_baseline"] = dict(base, note="the FP64 oracle, as c5_fp64 (timed once)")
    return res


def run_adaptive(n, steps, warmup, no_cpu):
    """advance_adaptive on the 4096² curvilinear upwind9 block (periodic,
    vortex + noise) with every step accepted (maxerr = inf): up to 40 embedded
    Bogacki-Shampine steps from a guess of 0.2/4095, the loop time growing by
    twice the step and the guess by 1.05 per step (after 40 steps the step is
    7x the first, a CFL number of about 0.4; many more would pass the
    stability limit)."""
    import torch
    from common import curvilinear

    X, Y, s, q = c2_state(n)
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

    steps, warmup = min(steps, 40), min(warmup, 2)
    s.advance_adaptive(span(warmup), g0, np.inf)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = s.advance_adaptive(span(steps), g0, np.inf)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    s.close()
    assert abs(g - g0 * 1.05 ** steps) <= 1e-12 * g
    pts = n * n
    ms = wall * 1e3 / steps
    unit = "Gpoint-steps/s"
    # the roofline of one accepted step taken as a whole (four stage launches,
    # the per-block error sums and their read-back)
    out = {"metric": "Gpoint-embedded-RK-steps/s (advance_adaptive, Bogacki-Shampine)", "value": pts * steps / wall / 1e9,
           "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"advance_adaptive on the {n}x{n} curvilinear upwind9 block, periodic, all steps accepted",
                      "nx": n, "ny": n, "operator": "upwind9"},
           "roofline": roofline("fp64", pts, ms / 4.0, bytes_per_point_step=BYTES_PER_POINT_ADAPTIVE_STEP,
                                definition="512 B per point per trial step of the fused embedded step "
                                           "(four stage launches; the embedded solution is never stored, its "
                                           "H-norm error is summed per tile), over the step's wall time")}
    if not no_cpu:
        m = 512
        Xs, Ys = curvilinear(m, m)

        def cpu(inner):
            from oracle import oracle as O
            ref = O.System([(Xs, Ys)], [(O.UPWIND9, O.UPWIND9)])
            ref.set_rhs_kind(O.RHS_UPWIND)
            ref.vortex(0.0, O.VortexParams.make([(-1.0, 0.0, 0.5, 1.0)], 0.5))
            t0 = time.perf_counter()
            ref.advance_adaptive(4.0 * 0.2 / (m - 1), 0.2 / (m - 1), np.inf)
            return m * m * 2, time.perf_counter() - t0

        w = Workload(None, None, None, None, None, cpu,
                     f"advance_adaptive, 2 accepted steps of a {m}x{m} block, oracle, 1 thread", 1)
        out["cpu_baseline"] = cpu_baseline(w, unit=unit)
    return out


def run_workload(name, precision, steps, warmup, tile, no_cpu, n=4096, schedule=-1):
    """The JSON object of one workload (its builder, timing, roofline and
    CPU baseline)."""
    if name == "adaptive":
        return run_adaptive(n, steps, warmup, no_cpu)
    if name == "c1":
        w, key = build_c1(precision), None
    elif name == "c2":
        w, key = build_c2(n, precision), "c2"
    elif name == "c3":
        w, key = build_c3(precision), "c3"
    elif name == "interp":
        w, key = build_interp(precision), None
    else:
        w, key = build_c5(precision)[0], "c5"
    t = time_steps(w, steps, warmup, tile, schedule)
    out = line(name, w, precision, t, steps, warmup, traffic_key=key)
    w.sys.close()
    if not no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    return out


def other_workloads(no_cpu):
    """Every configuration besides the headline block, each with its own
    timing, roofline and CPU baseline: (name, steps, warmup) per workload."""
    res = {}
    for name, steps, warmup in (("c1", 100, 10), ("c3", 500, 10), ("interp", 50, 5)):
        print(f"# {name}", file=sys.stderr, flush=True)
        res[name] = run_workload(name, "fp64", steps, warmup, -1, no_cpu)
    print("# c5", file=sys.stderr, flush=True)
    c5 = run_c5_pair(300, 10, no_cpu)
    res["c5_fp64"], res["c5_mixed"] = c5["fp64"], c5["mixed"]
    print("# adaptive", file=sys.stderr, flush=True)
    res["adaptive"] = run_adaptive(4096, 40, 2, no_cpu)
    return res


def calibrate_copy():
    """A plain 2 GiB device copy (the reference kernel for HBM byte
    counters), with the rates of a copy, a sum and a 3-read/1-write kernel."""
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
    return {"copy_bytes_read": n, "copy_bytes_written": n, "copy_gbs": rate(lambda: b.copy_(a), 2 * n),
            "sum_gbs": rate(lambda: a.sum(), n), "addcmul_3r1w_gbs": rate(lambda: torch.addcmul(c, d, e, out=o), 4 * q * 8)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c5", "interp", "adaptive"], default=None,
                    help="print the line of one workload only: c1 101^2 SBP8 far-field; c2 the 4096^2 curvilinear "
                         "block; c3 the 24-block Grid layout; c5 the 256-solve batch; interp a 2048^2 block under a "
                         "4095^2 block with 2:1 Interpolate interfaces; adaptive advance_adaptive on the 4096^2 block")
    ap.add_argument("--main-only", action="store_true", help="headline line without the other workloads")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="headline only: skip the mixed-form leg (throughput and 1000-step error vs FP64)")
    ap.add_argument("--tile", type=int, default=-1, help="force a tile shape index (-1: autotune)")
    ap.add_argument("--schedule", type=int, default=-1, choices=[-1, 0, 1, 2],
                    help="force the stage schedule (0 split over two streams, 1 merged, 2 merged after the "
                         "edge-SAT kernel; -1: autotune)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 2 GiB device copy (reference kernel for HBM byte counters)")
    ap.add_argument("--counters", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="summarize rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this bench (with --tile and "
                         "--workload) into HBM bytes per stage launch, calibrated by two runs of --calibrate-copy")
    args = ap.parse_args()
    if args.counters:
        from summationbyparts_amd import euler
        euler.set_gamma(1.4)
        wl = args.workload or "c2"
        pts = {"c2": args.n * args.n}.get(wl)
        if pts is None:
            from common import c3_layout, c5_batch
            pts = (sum(g[0].size for g in c3_layout(seed=3)[0]) if wl == "c3"
                   else sum(b["nx"] * b["ny"] for b in c5_batch()))
        print(json.dumps(summarize_counters(*args.counters, wl, args.tile, args.precision, pts)))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        print(json.dumps(calibrate_copy()))
        return
    from summationbyparts_amd import euler

    euler.set_gamma(1.4)
    if args.workload not in (None, "c2"):
        out = run_workload(args.workload, args.precision, args.steps, args.warmup, args.tile, args.no_cpu_baseline,
                           schedule=args.schedule)
        print(json.dumps(out), flush=True)
        return
    w = build_c2(args.n, args.precision)
    t = time_steps(w, args.steps, args.warmup, args.tile, args.schedule)
    # host -> device upload of one state through the boundary (not part of
    # the timed region: a run pays it once)
    q = w.sys.field()
    t1 = time.perf_counter()
    w.sys.set_state(0, q)
    torch.cuda.synchronize()
    upload_ms = (time.perf_counter() - t1) * 1e3
    out = line("c2", w, args.precision, t, args.steps, args.warmup, traffic_key="c2")
    out["config"]["state_upload_ms"] = upload_ms
    w.sys.close()
    if args.precision == "fp64" and args.tile < 0 and not args.no_mixed:
        print("# c2 mixed", file=sys.stderr, flush=True)
        out["mixed"] = mixed_leg(args.n, args.steps, args.warmup, t["value"])
    if not args.no_cpu_baseline:
        print("# c2 cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w)
    if args.workload is None and not args.main_only and args.n == 4096:
        out["workloads"] = other_workloads(args.no_cpu_baseline)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
