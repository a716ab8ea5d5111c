"""HDF5 output in the reference's layout (multigrid/src/system.rs:28-53,
87-117, 591-616), read back by the HDF5 library itself: the image's conda
Python (h5py on libhdf5 1.10) opens the files this writer makes, and its
layout, filters, chunking, links and values are compared with what the
reference's BaseSystem::new / System::output create. Skipped where that
Python is absent."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from summationbyparts_amd.hdf5 import Output  # noqa: E402

from common import have_h5py, read_h5  # noqa: E402

pytestmark = pytest.mark.skipif(not have_h5py(), reason="no Python with h5py (libhdf5) in this image")


def _grids(shapes, seed):
    rng = np.random.default_rng(seed)
    return [(rng.random((ny, nx)), rng.random((ny, nx))) for ny, nx in shapes]


def test_layout_and_values(tmp_path):
    shapes = [(5, 7), (3, 4)]
    grids = _grids(shapes, 1)
    rng = np.random.default_rng(2)
    fields = [[rng.standard_normal((4,) + s) for s in shapes] for _ in range(3)]
    path = tmp_path / "out.h5"
    with Output(path, ["grid0", "b"], grids) as o:
        for k, f in enumerate(fields):
            o.output(10 * k, f)
    info, arr = read_h5(path, tmp_path)
    assert sorted(info["groups"]) == ["b", "grid0"]
    # /t: u64, unlimited, chunk (1,), step indices; each group links to it
    t = info["datasets"]["t"]
    assert t["dtype"] == "uint64" and t["shape"] == [3] and t["maxshape"] == [None] and t["chunks"] == [1]
    assert t["compression"] is None and not t["shuffle"]
    assert arr["t"].tolist() == [0, 10, 20]
    assert info["links"] == {"grid0/t": "/t", "b/t": "/t"}
    assert info["grid_t"] == {"grid0": [0, 10, 20], "b": [0, 10, 20]}
    for name, (ny, nx), (x, y) in zip(["grid0", "b"], shapes, grids):
        for dim, v in (("x", x), ("y", y)):
            d = info["datasets"][f"{name}/{dim}"]
            assert d["dtype"] == "float64" and d["shape"] == [ny, nx] and d["maxshape"] == [ny, nx]
            assert d["chunks"] == [ny, nx] and d["compression"] == "gzip" and d["compression_opts"] == 9
            assert not d["shuffle"]
            assert np.array_equal(arr[f"{name}/{dim}"], v)
        for c, var in enumerate(("rho", "rhou", "rhov", "e")):
            d = info["datasets"][f"{name}/{var}"]
            assert d["dtype"] == "float64" and d["shape"] == [3, ny, nx] and d["maxshape"] == [None, ny, nx]
            assert d["chunks"] == [1, ny, nx] and d["compression"] == "gzip" and d["compression_opts"] == 3
            assert d["shuffle"]
            gi = ["grid0", "b"].index(name)
            assert np.array_equal(arr[f"{name}/{var}"], np.stack([f[gi][c] for f in fields]))


def test_empty_run_file(tmp_path):
    """The file right after set-up (BaseSystem::new): /t and the variables
    exist with no time slice."""
    path = tmp_path / "empty.h5"
    o = Output(path, ["a"], _grids([(4, 6)], 3))
    o.close()
    info, arr = read_h5(path, tmp_path)
    assert info["datasets"]["t"]["shape"] == [0]
    assert info["datasets"]["a/rho"]["shape"] == [0, 4, 6]
    assert arr["a/rho"].shape == (0, 4, 6)


def test_many_outputs_and_grids(tmp_path):
    """More time slices than one chunk-index node holds (64) and more grids
    than one group-index node reaches (32 symbol-table nodes of 8 entries):
    the writer builds multi-level indices."""
    ngrids, nout = 300, 70
    shapes = [(3 + i % 3, 2 + i % 4) for i in range(ngrids)]
    names = [f"block{i:03d}" for i in range(ngrids)]
    grids = _grids(shapes, 4)
    rng = np.random.default_rng(5)
    path = tmp_path / "many.h5"
    last = None
    with Output(path, names, grids) as o:
        for k in range(nout):
            last = [rng.standard_normal((4,) + s) for s in shapes]
            if k == 0:
                first = last
            o.output(k, last)
    info, arr = read_h5(path, tmp_path)
    assert sorted(info["groups"]) == names
    assert arr["t"].tolist() == list(range(nout))
    for i in (0, 7, 8, 150, 255, 256, 299):
        n = names[i]
        assert np.array_equal(arr[f"{n}/x"], grids[i][0])
        for c, var in enumerate(("rho", "rhou", "rhov", "e")):
            a = arr[f"{n}/{var}"]
            assert a.shape == (nout,) + shapes[i]
            assert np.array_equal(a[0], first[i][c]) and np.array_equal(a[-1], last[i][c])


def test_reopened_after_each_output(tmp_path):
    """The file is complete after every output (System::output leaves a
    readable file behind), and later outputs extend it."""
    path = tmp_path / "inc.h5"
    rng = np.random.default_rng(6)
    with Output(path, ["g"], _grids([(4, 5)], 7)) as o:
        for k in range(3):
            q = rng.standard_normal((4, 4, 5))
            o.output(k, [q])
            info, arr = read_h5(path, tmp_path)
            assert arr["t"].tolist() == list(range(k + 1))
            assert np.array_equal(arr["g/e"][k], q[3])
