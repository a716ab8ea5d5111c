"""Writes c3_oracle_error.json: the oracle's analytic vortex error
(System::error, multigrid/src/system.rs:618-635, the sum of the blocks'
Field::h2_err) of the C3 layout (24 Cartesian blocks, sizes drawn from seed 3
in [64, 1024], Grid(j) interfaces, upwind4, vortex initial state) after 500
RK4 steps of dt = 0.2·min over blocks of min(1/(nx-1), 1/(ny-1)), one
thread per block. The GPU test compares its own 500-step error with it.

Run from the repository root: python tests/golden/make_c3_error.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from common import VORTEX, MACH, c3_dt, c3_layout  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main(steps=500, threads=8):
    grids, nbs = c3_layout(seed=3)
    ref = O.System(grids, [(O.UPWIND4, O.UPWIND4)] * len(grids))
    for i, nb in enumerate(nbs):
        for d, j in nb.items():
            ref.set_bc(i, d, O.BC_GRID, j)
    vo = O.VortexParams.make(list(VORTEX), MACH)
    ref.vortex(0.0, vo)
    dt = c3_dt(grids)
    t0 = time.perf_counter()
    ref.advance(dt, steps, threads)
    out = {"steps": steps, "dt": dt, "time": ref.time, "error": ref.error_vortex(vo),
           "blocks": len(grids), "points": int(sum(g[0].size for g in grids)),
           "seconds": time.perf_counter() - t0}
    with open(os.path.join(HERE, "c3_oracle_error.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
