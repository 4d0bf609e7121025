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


def hull_triangles(points):
    """The outward triangles (counter-clockwise seen from outside) of the convex hull of the
    points, indexing the input; None if the cloud is flat.  The contact-manifold generators build
    a hull's faces from these."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    n = C.c_size_t()
    L = _lib.load()
    _lib.check(L.ncx_hull_triangulate(_lib.ptr(pts), len(pts), None, 0, C.byref(n)))
    if n.value == 0:
        return None
    out = np.zeros((n.value, 3), np.uint32)
    _lib.check(L.ncx_hull_triangulate(_lib.ptr(pts), len(pts), _lib.ptr(out), n.value, C.byref(n)))
    return out


def hull_polyhedron(points, tris):
    """ConvexHull::try_new's polygonal faces, edges and adjacency (f32) from a hull's outward
    triangles; None where the reference returns None."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    tr = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    n, nt = len(pts), len(tr)
    fn = np.zeros((nt, 3), np.float32)
    ff, fc = np.zeros(nt, np.uint32), np.zeros(nt, np.uint32)
    lv, le = np.zeros(3 * nt, np.uint32), np.zeros(3 * nt, np.uint32)
    ev, ed = np.zeros((3 * nt, 2), np.uint32), np.zeros((3 * nt, 3), np.float32)
    vf, vc = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    af, ae = np.zeros(3 * nt, np.uint32), np.zeros(3 * nt, np.uint32)
    nf, nl, ne = C.c_size_t(), C.c_size_t(), C.c_size_t()
    p = _lib.ptr
    rc = _lib.load().ncx_hull_polyhedron(p(pts), n, p(tr), nt, p(fn), p(ff), p(fc), C.byref(nf), p(lv), p(le),
                                         C.byref(nl), p(ev), p(ed), C.byref(ne), p(vf), p(vc), p(af), p(ae))
    if rc != 0:
        return None
    nf, nl, ne = nf.value, nl.value, ne.value
    na = int(vc.sum())
    return dict(face_normals=fn[:nf], face_first=ff[:nf], face_count=fc[:nf], loop_vertex=lv[:nl], loop_edge=le[:nl],
                edge_vertices=ev[:ne], edge_dirs=ed[:ne], vertex_first=vf, vertex_count=vc, adj_face=af[:na],
                adj_edge=ae[:na])


def polyhedra_ok(pool, offsets, counts, nthreads=None):
    """Per hull of a vertex pool: whether ConvexHull::try_new succeeds on its triangles (the
    reference has no such ConvexHull where it does not)."""
    pts = np.ascontiguousarray(pool, np.float32).reshape(-1, 3)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    counts = np.ascontiguousarray(counts, np.uint32)
    ok = np.zeros(len(counts), np.uint8)
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    _lib.check(_lib.load().ncx_hull_polyhedra_check(_lib.ptr(pts), _lib.ptr(offsets), _lib.ptr(counts), len(counts),
                                                    _lib.ptr(ok), int(nthreads)))
    return ok.astype(bool)
