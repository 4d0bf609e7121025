"""CPU checks of the drop-in boundary: libncx.so loads and exports every symbol that
include/ncx.h declares; host-only entry points work without a GPU."""
import os
import re

import numpy as np
import pytest

from ncollide_amd import _lib, hull, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "ncx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ncx_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    names = declared_symbols()
    missing = [n for n in names if n not in _lib.SIGNATURES]
    assert not missing, missing


def test_init_without_gpu_fails_loudly():
    import ctypes as C
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.ncx_init(C.byref(h), 0, None) == -4  # NCX_E_NODEVICE
    assert b"no HIP device" in lib.ncx_last_error()


def test_context_refuses_cpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import ncollide_amd
    with pytest.raises(ncollide_amd.NcxError):
        ncollide_amd.Context(0)


def test_convex_hull_matches_qhull():
    from scipy.spatial import ConvexHull
    rng = np.random.default_rng(0)
    for _ in range(100):
        n = int(rng.integers(8, 257))
        x = (workloads.unit_vectors(rng, n) * rng.uniform(0.2, 1, 3) * (1 + rng.uniform(-0.02, 0.02, (n, 1))))
        x = x.astype(np.float32)
        assert np.array_equal(hull.convex_hull_vertices(x), np.sort(ConvexHull(x.astype(np.float64)).vertices))


def test_convex_hull_degenerate():
    flat = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0.5, 0.5, 0]], np.float32)
    assert hull.convex_hull_vertices(flat) is None
    cube = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)] + [[0.5, 0.5, 0.5]], np.float32)
    assert np.array_equal(hull.convex_hull_vertices(cube), np.arange(8))


def test_convex_hull_batch_equals_single():
    rng = np.random.default_rng(1)
    pts, offs, cnts = workloads.ellipsoid_clouds(rng, 50)
    ids, hc = hull.convex_hull_batch(pts, offs, cnts, nthreads=4)
    for i in range(50):
        single = hull.convex_hull_vertices(pts[offs[i]:offs[i] + cnts[i]])
        assert np.array_equal(ids[offs[i]:offs[i] + hc[i]], single)


def test_workload_shapes():
    w = workloads.plain_pipeline()
    assert len(w["records"]) == 1024 and w["poses"].shape == (1024, 8)
    assert np.allclose(np.linalg.norm(w["poses"][:, 4:8], axis=1), 1, atol=1e-6)
    pts, tri = workloads.heightfield_mesh()
    assert pts.shape == (257 * 257, 3) and tri.shape == (131072, 3)
    # every triangle is CCW seen from +z
    a, b, c = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    assert np.all(np.cross(b - a, c - a)[:, 2] > 0)
    o, d = workloads.heightfield_rays(n=1000)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6) and np.all(d[:, 2] < 0)
