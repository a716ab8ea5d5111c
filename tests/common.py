"""Shared inputs of the tests: grids, vortex parameters and seeded noise of
the benchmark workloads, and comparison helpers."""
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


def rel_global(a, b):
    """max|a-b| / max|b| over all components. Used for right-hand sides,
    whose components cancel to very different magnitudes (the density
    component of a trad RHS is a difference of O(idx) fluxes)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
