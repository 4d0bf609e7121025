"""One collision step of a whole scene (broad phase + contacts), capturable in a hipGraph.

Python face of ncx_world (include/ncx.h): the hot path of CollisionWorld::update
(src/pipeline/world.rs:104-119) for a ShapeSet whose shapes are all broad-phase proxies.
"""
import ctypes as C

import torch

from . import _lib


class World:
    def __init__(self, shapes, margin, prediction=0.0, max_pairs=1 << 20, rebuild_every=2):
        self.ctx = shapes.ctx
        self.shapes = shapes
        h = C.c_void_p()
        _lib.check(self.ctx.lib.ncx_world_create(self.ctx.h, shapes.h, float(margin), float(prediction),
                                                 int(max_pairs), C.byref(h)))
        self.h = h
        self.max_pairs = int(max_pairs)
        _lib.check(self.ctx.lib.ncx_world_set_rebuild_every(self.h, int(rebuild_every)))
        self._mlp = None

    def set_warmstart(self, mlp):
        self._mlp = mlp
        _lib.check(self.ctx.lib.ncx_world_set_warmstart(self.h, mlp.h if mlp is not None else None))

    def set_contacts(self, enable):
        """Runs each step's narrow phase (default) or stops the step at the pair list."""
        _lib.check(self.ctx.lib.ncx_world_set_contacts(self.h, 1 if enable else 0))

    def query_stats(self):
        """Internal nodes visited by one pair-count traversal of the current tree (measurement)."""
        v = C.c_ulonglong()
        _lib.check(self.ctx.lib.ncx_world_query_stats(self.h, C.byref(v)))
        return v.value

    def step(self, poses, use_graph=True):
        """poses: (n, 8) float32 CUDA tensor (keep the same tensor across steps to replay the graph)."""
        _lib.check(self.ctx.lib.ncx_world_step(self.h, _lib.ptr(poses), 1 if use_graph else 0))

    def sync(self):
        n = C.c_size_t()
        ovf = C.c_size_t()
        _lib.check(self.ctx.lib.ncx_world_sync(self.h, C.byref(n), C.byref(ovf)))
        return n.value, ovf.value

    def results(self):
        """Copies of the last step's outputs: dict of CUDA tensors (ia, ib, has_contact, world1, world2,
        normal, depth, gjk_iters) over the current pairs."""
        n, ovf = self.sync()
        ptrs = [C.c_void_p() for _ in range(8)]
        _lib.check(self.ctx.lib.ncx_world_outputs(self.h, *[C.byref(p) for p in ptrs]))
        dev = self.ctx.device
        out = {}
        spec = [("ia", torch.int32, 1), ("ib", torch.int32, 1), ("has_contact", torch.uint8, 1),
                ("world1", torch.float32, 3), ("world2", torch.float32, 3), ("normal", torch.float32, 3),
                ("depth", torch.float32, 1), ("gjk_iters", torch.int32, 1)]
        for (name, dt, width), p in zip(spec, ptrs):
            t = torch.empty((max(n, 1), width) if width > 1 else max(n, 1), dtype=dt, device=dev)
            esz = t.element_size() * width
            if n:
                _memcpy_d2d(t.data_ptr(), p.value, n * esz)
            out[name] = t[:n]
        out["has_contact"] = out["has_contact"].bool()
        out["epa_overflow"] = ovf
        return out

    def stored_aabbs(self):
        p = C.c_void_p()
        _lib.check(self.ctx.lib.ncx_world_stored_aabbs(self.h, C.byref(p)))
        n = self.shapes.n
        t = torch.empty((6, n), dtype=torch.float32, device=self.ctx.device)
        _memcpy_d2d(t.data_ptr(), p.value, 24 * n)
        return t

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.ctx.lib.ncx_world_destroy(self.h)
        except Exception:
            pass


_hip = None


def _memcpy_d2d(dst, src, nbytes):
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    torch.cuda.synchronize()
    rc = _hip.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), 3)
    if rc != 0:
        raise _lib.NcxError(f"hipMemcpy failed ({rc})")
