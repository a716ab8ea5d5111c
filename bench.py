"""Benchmark: Euler RK4 step throughput on one MI355X.

Workload (default): one curvilinear 4096x4096 block, upwind9 with upwind
dissipation, isentropic vortex (centre (-1, 0), r* 0.5, eps 1, Mach 0.5) plus
seeded +-1e-3 density noise, periodic SATs, dt = 0.2/4095, FP64. A step is one
classical RK4 step (four fused stage kernels) over the whole block with the
state resident in HBM.

Prints one JSON line: throughput in Gpoint-RK4-steps/s, the HBM roofline of
the stage kernel (algorithmic bytes per launch / mean launch duration from HIP
events), and the CPU oracle timed on this host on a bounded sample.
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
# Algorithmic bytes per point per RK4 step of the fused schedule.
# FP64: stage 1 reads y_n and 5 metric planes, writes acc and Y2 (136 B);
# stages 2-3 read Y_s, metrics, y_n, acc and write acc, Y_{s+1} (200 B each);
# stage 4 reads Y4, metrics, acc and writes y_{n+1} (136 B): 672 B.
# Mixed: the stage inputs are y_n + Δ_s with FP32 increments Δ_s: stage 1
# reads y_n, metrics, writes acc, Δ_1 (120 B); stages 2-3 read y_n, Δ_s,
# metrics, acc and write acc, Δ_{s+1} (168 B each); stage 4 reads y_n, Δ_3,
# metrics, acc and writes y_{n+1} (152 B): 608 B.
BYTES_PER_POINT_STEP = {"fp64": 672.0, "mixed": 608.0}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stage_kernel_upwind9_4096_traffic.json")


def build_c2(n, precision):
    from common import curvilinear, density_noise
    from summationbyparts_amd import _lib, euler, operators

    euler.set_gamma(1.4)
    X, Y = curvilinear(n, n)
    sys_ = euler.System(X, Y, operators.Upwind9())
    if precision == "mixed":
        sys_.set_precision(_lib.MIXED)
    sys_.vortex(0.0, euler.VortexParameters([euler.Vortice(-1.0, 0.0, 0.5, 1.0)], 0.5))
    q = sys_.field()
    q[0] += density_noise(1, 0, n, n)
    sys_.set_state(0, q)
    sys_.set_rhs_kind(_lib.RHS_UPWIND)
    return sys_, X, Y, q


def cpu_baseline(X, Y, q, steps):
    """The oracle (single thread, the reference's one-thread-per-block
    semantics for one block) on `steps` RK4 steps of the same block."""
    from oracle import oracle as O

    n = X.shape[0]
    ref = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
    ref.set_state(0, q)
    dt = 0.2 / (n - 1)
    t0 = time.perf_counter()
    ref.advance(dt, steps, 1)
    sec = time.perf_counter() - t0
    return {"value": n * n * steps / sec / 1e9, "unit": UNIT, "cores": 1, "kind": "port",
            "sample": f"{steps} RK4 steps of the {n}x{n} upwind9 curvilinear block, oracle (C++ restatement), "
                      f"1 thread, {os.cpu_count()} host cores visible"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile", type=int, default=-1, help="force a tile shape index (-1: autotune)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 2 GiB device copy (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        a = torch.ones(1 << 28, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    n = args.n
    sys_, X, Y, q0 = build_c2(n, args.precision)
    if args.tile >= 0:
        sys_.set_tile_shape(args.tile)
    dt = 0.2 / (n - 1)
    sys_.advance_n(dt, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sys_.advance_steps(dt, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # host -> device upload of one state through the boundary (not part of
    # the timed region: a run pays it once)
    t1 = time.perf_counter()
    sys_.set_state(0, q0)
    torch.cuda.synchronize()
    upload_ms = (time.perf_counter() - t1) * 1e3
    # after the timed region: HIP-event time of 8 graph-replayed steps, then
    # the mean duration of one stage-kernel launch over 8 more steps (events
    # on the library's stream around each launch)
    ms_events, ms_stage = sys_.profile(dt, 8)
    tile = sys_.tile_shape()
    pts = n * n
    value = pts * args.steps / wall / 1e9
    bytes_per_launch = BYTES_PER_POINT_STEP[args.precision] / 4.0 * pts
    achieved = bytes_per_launch / (ms_stage * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        if tr.get("n") == n and tr.get("precision", "fp64") == args.precision:
            traffic = tr.get("hbm_bytes_per_launch")
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
        "config": {"workload": f"single curvilinear block {n}x{n}, upwind9 + dissipation, vortex + 1e-3 density "
                               f"noise, periodic SATs, RK4 dt=0.2/{n - 1}", "nx": n, "ny": n,
                   "operator": "upwind9", "precision": args.precision, "tile": list(tile),
                   "event_ms_per_step": ms_events / 8, "stage_launch_ms": ms_stage,
                   "state_upload_ms": upload_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(X, Y, q0, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
