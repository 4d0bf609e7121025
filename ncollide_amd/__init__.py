"""ncollide_amd -- ncollide3d's collision hot path on one MI355X (gfx950).

The package mirrors the reference's API surface for that path:
  DBVTBroadPhase (pipeline::broad_phase), query.contact / query.distance
  (query::contact, query::distance) and TriMesh ray casting (RayCast for TriMesh),
batched over device-resident arrays and executed by the hand-written HIP kernels
of libncx.so (C ABI in include/ncx.h).  There is no CPU fallback: every query
raises if the library or the GPU is missing.
"""
from ._lib import NcxError, load as load_library
from .context import Context, default_context
from .shape import Ball, Cuboid, Capsule, ConvexHull, ShapeSet, BALL, CUBOID, CAPSULE, CONVEX_HULL
from .broad_phase import DBVTBroadPhase
from .ray import TriMesh, NO_HIT
from .manifold import ContactManifolds
from . import query, workloads, math, manifold

__all__ = ["NcxError", "load_library", "Context", "default_context", "Ball", "Cuboid", "Capsule", "ConvexHull",
           "ShapeSet", "DBVTBroadPhase", "TriMesh", "NO_HIT", "query", "workloads", "math",
           "BALL", "CUBOID", "CAPSULE", "CONVEX_HULL"]
