"""Shapes and shape sets (the hot path's Ball, Cuboid, Capsule, ConvexHull).

Mirrors src/shape/{ball,cuboid,capsule,convex}.rs.  A ShapeSet packs a list of
shapes into the device layout of include/ncx.h (ncx_shape records plus a hull
vertex pool in three float arrays).
"""
import ctypes as C

import numpy as np

from . import _lib
from .context import default_context

BALL, CUBOID, CAPSULE, CONVEX_HULL = 0, 1, 2, 3

SHAPE_DTYPE = np.dtype([("type", "<u4"), ("hull_offset", "<u4"), ("hull_count", "<u4"),
                        ("p0", "<f4"), ("p1", "<f4"), ("p2", "<f4"), ("pad", "<u4", (2,))])


class Ball:
    """src/shape/ball.rs: a ball of the given radius centred at the origin."""

    def __init__(self, radius):
        self.radius = float(radius)


class Cuboid:
    """src/shape/cuboid.rs: a box with the given half extents."""

    def __init__(self, half_extents):
        self.half_extents = np.asarray(half_extents, np.float32).reshape(3)


class Capsule:
    """src/shape/capsule.rs: segment along local y of half height `half_height`, dilated by `radius`."""

    def __init__(self, half_height, radius):
        self.half_height = float(half_height)
        self.radius = float(radius)


class ConvexHull:
    """src/shape/convex.rs: the support function scans `points` (local frame).

    `try_from_points` keeps the input points that are vertices of the convex
    hull (in input order), see ncollide_amd.hull.
    """

    def __init__(self, points):
        self.points = np.ascontiguousarray(points, np.float32).reshape(-1, 3)

    @staticmethod
    def try_from_points(points):
        from .hull import convex_hull_vertices
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        ids = convex_hull_vertices(pts)
        if ids is None or len(ids) < 4:
            return None
        return ConvexHull(pts[ids])


def pack(shapes):
    """Packs shape objects into (records, hull pool xyz)."""
    rec = np.zeros(len(shapes), SHAPE_DTYPE)
    pools = []
    off = 0
    for i, s in enumerate(shapes):
        if isinstance(s, Ball):
            rec[i]["type"] = BALL
            rec[i]["p0"] = s.radius
        elif isinstance(s, Cuboid):
            rec[i]["type"] = CUBOID
            rec[i]["p0"], rec[i]["p1"], rec[i]["p2"] = s.half_extents
        elif isinstance(s, Capsule):
            rec[i]["type"] = CAPSULE
            rec[i]["p0"], rec[i]["p1"] = s.half_height, s.radius
        elif isinstance(s, ConvexHull):
            rec[i]["type"] = CONVEX_HULL
            rec[i]["hull_offset"] = off
            rec[i]["hull_count"] = len(s.points)
            pools.append(s.points)
            off += len(s.points)
        else:
            raise TypeError(f"unsupported shape {type(s).__name__}")
    pool = np.concatenate(pools) if pools else np.zeros((0, 3), np.float32)
    return rec, np.ascontiguousarray(pool, np.float32)


class ShapeSet:
    """A device-resident set of shapes, indexed 0..n-1 by the batched queries."""

    def __init__(self, shapes=None, records=None, pool=None, ctx=None):
        self.ctx = ctx or default_context()
        if records is None:
            records, pool = pack(shapes)
        self.records = np.ascontiguousarray(records, SHAPE_DTYPE)
        self.pool = np.ascontiguousarray(pool, np.float32).reshape(-1, 3)
        px = np.ascontiguousarray(self.pool[:, 0])
        py = np.ascontiguousarray(self.pool[:, 1])
        pz = np.ascontiguousarray(self.pool[:, 2])
        h = C.c_void_p()
        _lib.check(self.ctx.lib.ncx_shapes_create(self.ctx.h, _lib.ptr(self.records), len(self.records),
                                                  _lib.ptr(px), _lib.ptr(py), _lib.ptr(pz), len(self.pool),
                                                  C.byref(h)))
        self.h = h
        self.n = len(self.records)

    @property
    def types(self):
        return self.records["type"]

    # first-tier variants: (lanes per pair, EPA arena vertices)
    VARIANTS = ((8, 44), (16, 44), (32, 44), (8, 24), (16, 24))
    # distance queries have no EPA arena: lanes per pair of each variant
    DISTANCE_VARIANTS = ((8, 0), (16, 0), (32, 0), (4, 0), (2, 0))

    def tuning(self, query="contact"):
        """The autotuner's choice per vertex-count bucket (no hull, <= 32, <= 96, more vertices):
        the variant (lanes per pair, first-tier EPA arena vertices, 0 for distance) or None before
        tuning, and each variant's sample time (ms; 0 for variants not tried)."""
        table = self.VARIANTS if query == "contact" else self.DISTANCE_VARIANTS
        nv = len(table)
        choice = (C.c_int32 * 4)()
        ms = (C.c_float * (4 * nv))()
        _lib.check(self.ctx.lib.ncx_shapes_tuning(self.h, 0 if query == "contact" else 1, choice, ms))
        return ([table[c] if c >= 0 else None for c in choice], [list(ms[nv * b:nv * b + nv]) for b in range(4)])

    def set_tuning(self, variants, query="contact"):
        """Pins the first-tier variant of each bucket: a list of 4 indices into VARIANTS (None =
        tune on the next large batch).  Results do not depend on the variant."""
        choice = (C.c_int32 * 4)(*[-1 if v is None else int(v) for v in variants])
        _lib.check(self.ctx.lib.ncx_shapes_set_tuning(self.h, 0 if query == "contact" else 1, choice))

    def params(self):
        return np.stack([self.records["p0"], self.records["p1"], self.records["p2"]], 1)

    def aabbs(self, poses, margin=0.0):
        """World AABBs (Shape::aabb, loosened by margin) as a (6, n) float32 CUDA tensor."""
        import torch
        out = torch.empty((6, self.n), dtype=torch.float32, device=self.ctx.device)
        _lib.check(self.ctx.lib.ncx_shapes_aabbs_dev(self.h, _lib.ptr(poses), float(margin),
                                                     _lib.ptr6([out[d] for d in range(6)])))
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.ctx.lib.ncx_shapes_destroy(self.h)
        except Exception:
            pass
