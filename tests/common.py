"""Shared inputs of the tests: grids, vortex parameters and seeded noise of
the benchmark workloads, and comparison helpers."""
import json
import os
import subprocess

import numpy as np

GAMMA = 1.4
# C1/C2 vortex: centre (-1, 0), rstar 0.5, eps 1.0, Mach 0.5 (euler/tests/convergence.rs:23-35)
VORTEX = ((-1.0, 0.0, 0.5, 1.0),)
MACH = 0.5


def cartesian(nx, ny, lo=-5.0, hi=5.0):
    """x = linspace broadcast along rows, y along columns (convergence.rs:12-21)."""
    x = np.linspace(lo, hi, nx)
    y = np.linspace(lo, hi, ny)
    return np.broadcast_to(x[None, :], (ny, nx)).copy(), np.broadcast_to(y[:, None], (ny, nx)).copy()


def curvilinear(nx, ny, amp=0.1, lo=-5.0, hi=5.0):
    """x = ξ + amp·sin(2πη)·cos(πξ), y = η + amp·sin(2πξ) on the unit square,
    scaled to [lo, hi]^2."""
    xi = np.arange(nx, dtype=np.float64) / (nx - 1)
    eta = np.arange(ny, dtype=np.float64) / (ny - 1)
    XI, ETA = np.meshgrid(xi, eta)
    x = XI + amp * np.sin(2 * np.pi * ETA) * np.cos(np.pi * XI)
    y = ETA + amp * np.sin(2 * np.pi * XI)
    return lo + (hi - lo) * x, lo + (hi - lo) * y


def splitmix64(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def density_noise(seed, block, ny, nx, amplitude=1e-3):
    """Uniform noise in [-amplitude, amplitude) keyed by (seed, block, j, i)
    through SplitMix64, so every consumer builds the same field."""
    with np.errstate(over="ignore"):
        j = np.arange(ny, dtype=np.uint64)[:, None]
        i = np.arange(nx, dtype=np.uint64)[None, :]
        key = (np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(block)) * np.uint64(1 << 40) \
            + j * np.uint64(1 << 20) + i
        r = splitmix64(key)
    u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return amplitude * (2.0 * u - 1.0)


def rel_maxnorm(a, b):
    """max|a-b| / max|b| per component, maximum over components."""
    a = np.asarray(a)
    b = np.asarray(b)
    return max(float(np.max(np.abs(a[c] - b[c])) / max(np.max(np.abs(b[c])), 1e-300)) for c in range(a.shape[0]))


def rel_maxnorm_blocks(a_blocks, b_blocks):
    """Multi-block rel_maxnorm: max|a - b| per component over all blocks,
    relative to that component's largest magnitude over all blocks (a block
    far from the vortex holds ρv of ~1e-70 to ~1e-16, so per-block scaling
    would compare rounding noise with itself)."""
    err = 0.0
    for c in range(4):
        scale = max(float(np.max(np.abs(b[c]))) for b in b_blocks)
        d = max(float(np.max(np.abs(a[c] - b[c]))) for a, b in zip(a_blocks, b_blocks))
        err = max(err, d / max(scale, 1e-300))
    return err


def rel_global(a, b):
    """max|a-b| / max|b| over all components. Used for right-hand sides,
    whose components cancel to very different magnitudes (the density
    component of a trad RHS is a difference of O(idx) fluxes)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def c3_layout(seed=3, lo=64, hi=1024, cols=6, rows=4, extent=10.0):
    """Batched multi-block workload C3: a cols x rows periodic layout of
    Cartesian blocks tiling [-extent, extent]^2. Column widths nx_c and row
    heights ny_r are drawn from `seed` in [lo, hi], so every shared edge has
    the same number of points on both sides (Grid(j) couplings). Returns
    (grids, neighbours) with neighbours[i] = {direction: block}."""
    rng = np.random.default_rng(seed)
    widths = rng.integers(lo, hi + 1, size=cols)
    heights = rng.integers(lo, hi + 1, size=rows)
    xs = np.linspace(-extent, extent, cols + 1)
    ys = np.linspace(-extent, extent, rows + 1)
    grids, index = [], {}
    for r in range(rows):
        for c in range(cols):
            nx, ny = int(widths[c]), int(heights[r])
            x = np.linspace(xs[c], xs[c + 1], nx)
            y = np.linspace(ys[r], ys[r + 1], ny)
            index[(r, c)] = len(grids)
            grids.append((np.broadcast_to(x[None, :], (ny, nx)).copy(), np.broadcast_to(y[:, None], (ny, nx)).copy()))
    # directions as in the C ABI: north, south, east, west
    nbs = []
    for r in range(rows):
        for c in range(cols):
            nbs.append({0: index[((r + 1) % rows, c)], 1: index[((r - 1) % rows, c)],
                        2: index[(r, (c + 1) % cols)], 3: index[(r, (c - 1) % cols)]})
    return grids, nbs


def c3_dt(grids):
    """0.2 * min over blocks of min(1/(nx-1), 1/(ny-1))."""
    return 0.2 * min(min(1.0 / (g[0].shape[1] - 1), 1.0 / (g[0].shape[0] - 1)) for g in grids)


def c5_batch(seed=5, count=256, lo=48, hi=512):
    """Batched independent solves C5: per solve (ny, nx) in [lo, hi]^2, vortex
    Mach in [0.3, 0.7], eps in [0.5, 2.0] and curvilinear amplitude in
    [0, 0.15], all drawn from `seed`; centre (-1, 0) on [-5, 5]^2 with
    r* = 1: with r* = 0.5 the strongest draws (eps 2, Mach 0.7) have negative
    density in the core (1 - eps²(γ-1)γM⁴/(8π²r*²)·e^(1/r*²) < 0), and r* = 1
    also keeps the core five cells wide on the coarsest grids.
    Returns a list of dicts (nx, ny, mach, eps, amp, rstar, dt) with the
    solve's step dt = 0.2 * min(1/(nx-1), 1/(ny-1))."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx, ny = (int(v) for v in rng.integers(lo, hi + 1, size=2))
        mach = float(rng.uniform(0.3, 0.7))
        eps = float(rng.uniform(0.5, 2.0))
        amp = float(rng.uniform(0.0, 0.15))
        out.append(dict(nx=nx, ny=ny, mach=mach, eps=eps, amp=amp, rstar=1.0,
                        dt=0.2 * min(1.0 / (nx - 1), 1.0 / (ny - 1))))
    return out


# ---------------------------------------------------------------------------
# HDF5 files read back by the HDF5 library: the image's conda Python carries
# h5py on libhdf5 1.10 (the system Python has no HDF5 module)
H5PY = "/opt/conda/bin/python3.9"


def have_h5py():
    if not os.path.exists(H5PY):
        return False
    r = subprocess.run([H5PY, "-c", "import h5py"], capture_output=True)
    return r.returncode == 0


# reads every dataset, link and storage property of a file with h5py
_READER = r"""
import h5py, json, sys, numpy as np
out = {"datasets": {}, "links": {}, "groups": []}
with h5py.File(sys.argv[1], "r") as f:
    def visit(g, prefix):
        for name in g:
            path = prefix + name
            link = g.get(name, getlink=True)
            if isinstance(link, h5py.SoftLink):
                out["links"][path] = link.path
                continue
            obj = g[name]
            if isinstance(obj, h5py.Group):
                out["groups"].append(path)
                visit(obj, path + "/")
            else:
                np.save(sys.argv[2] + "/" + path.replace("/", "__") + ".npy", obj[()])
                out["datasets"][path] = {
                    "dtype": str(obj.dtype), "shape": list(obj.shape),
                    "maxshape": [m for m in obj.maxshape], "chunks": list(obj.chunks) if obj.chunks else None,
                    "compression": obj.compression, "compression_opts": obj.compression_opts,
                    "shuffle": bool(obj.shuffle)}
    visit(f, "")
    out["grid_t"] = {g: [int(v) for v in f[g]["t"][()]] for g in out["groups"]}
print(json.dumps(out))
"""


def read_h5(path, tmp):
    d = tmp / "arrays"
    d.mkdir(exist_ok=True)
    r = subprocess.run([H5PY, "-c", _READER, str(path), str(d)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout)
    arrays = {p: np.load(d / (p.replace("/", "__") + ".npy")) for p in info["datasets"]}
    return info, arrays
