"""Convex hull vertex selection (ConvexHull::try_from_points input preparation), host C++ in libncx."""
import ctypes as C
import os

import numpy as np

from . import _lib


def convex_hull_vertices(points):
    """Indices (ascending) of the input points that are vertices of their convex hull; None if degenerate."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    out = np.zeros(max(len(pts), 1), np.uint32)
    n = C.c_size_t()
    _lib.check(_lib.load().ncx_convex_hull(_lib.ptr(pts), len(pts), _lib.ptr(out), C.byref(n)))
    if n.value == 0:
        return None
    return out[:n.value]


def convex_hull_batch(points, offsets, counts, nthreads=None):
    """Hull vertex ids of many point clouds at once.  Returns (ids pool, hull counts): cloud i's
    vertex ids are ids[offsets[i] : offsets[i] + hull_counts[i]] (indices into the cloud)."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    counts = np.ascontiguousarray(counts, np.uint32)
    ids = np.zeros(max(len(pts), 1), np.uint32)
    hc = np.zeros(len(counts), np.uint32)
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    _lib.check(_lib.load().ncx_convex_hull_batch(_lib.ptr(pts), _lib.ptr(offsets), _lib.ptr(counts), len(counts),
                                                 _lib.ptr(ids), _lib.ptr(hc), int(nthreads)))
    return ids, hc
