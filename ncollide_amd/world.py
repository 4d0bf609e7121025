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

    def spill_stats(self):
        """Of the last step's pair search: (leaves with more than 512 later partners, of those the
        leaves whose further partners went to the spill chunks rather than a second search)."""
        out = (C.c_uint32 * 2)()
        _lib.check(self.ctx.lib.ncx_world_spill_stats(self.h, out))
        return out[0], out[1]

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

    def set_temporal_warmstart(self, enable=True):
        """The polyhedron manifold generator's last_gjk_dir kept per pair across steps (the
        reference's own warm start of polytope pairs)."""
        _lib.check(self.ctx.lib.ncx_world_set_temporal_warmstart(self.h, 1 if enable else 0))

    def temporal_dirs(self):
        """(init_dir, result direction) per pair of the last step under the temporal warm start."""
        n, _ = self.sync()
        pi, pr = C.c_void_p(), C.c_void_p()
        _lib.check(self.ctx.lib.ncx_world_temporal_dirs(self.h, C.byref(pi), C.byref(pr)))
        out = []
        for p in (pi, pr):
            t = torch.empty((max(n, 1), 3), dtype=torch.float32, device=self.ctx.device)
            if n:
                _memcpy_d2d(t.data_ptr(), p.value, 12 * n)
            out.append(t[:n])
        return out[0], out[1]

    def warmstart_dirs(self):
        """The learned init_dir of each pair of the last step ((npairs, 3) CUDA tensor, zero rows = None)."""
        n, _ = self.sync()
        p = C.c_void_p()
        _lib.check(self.ctx.lib.ncx_world_warmstart_dirs(self.h, C.byref(p)))
        t = torch.empty((max(n, 1), 3), dtype=torch.float32, device=self.ctx.device)
        if n:
            _memcpy_d2d(t.data_ptr(), p.value, 12 * n)
        return t[:n]

    def stored_aabbs(self):
        p = C.c_void_p()
        _lib.check(self.ctx.lib.ncx_world_stored_aabbs(self.h, C.byref(p)))
        n = self.shapes.n
        t = torch.empty((6, n), dtype=torch.float32, device=self.ctx.device)
        _memcpy_d2d(t.data_ptr(), p.value, 24 * n)
        return t

    # -- world-level queries through the broad phase (after a step; handles are shape indices)
    def _rays(self, origins, dirs):
        return _lib.soa_columns(torch.cat([torch.as_tensor(origins, dtype=torch.float32).reshape(-1, 3).cpu(),
                                           torch.as_tensor(dirs, dtype=torch.float32).reshape(-1, 3).cpu()], 1),
                                6, self.ctx.device)

    def first_interference_with_rays(self, poses, origins, dirs, max_toi):
        """CollisionWorld::first_interference_with_ray per ray (solid casts): (handle, toi, normal)
        CUDA tensors; handle -1 (0xFFFFFFFF) and toi -1 on a miss.  poses: the last step's poses."""
        cols = self._rays(origins, dirs)
        n = cols[0].numel()
        dev = self.ctx.device
        h = torch.empty(n, dtype=torch.int32, device=dev)
        t = torch.empty(n, dtype=torch.float32, device=dev)
        nr = torch.empty((n, 3), dtype=torch.float32, device=dev)
        _lib.check(self.ctx.lib.ncx_world_first_interference_with_rays(self.h, _lib.ptr(poses), _lib.ptrs(cols), n,
                                                                       float(max_toi), _lib.ptr(h), _lib.ptr(t),
                                                                       _lib.ptr(nr)))
        return h, t, nr

    def interferences_with_rays(self, poses, origins, dirs, max_toi, aabb_only=False):
        """CollisionWorld::interferences_with_ray per ray: CSR (offsets (n+1,), handles, toi, normal)."""
        cols = self._rays(origins, dirs)
        n = cols[0].numel()
        dev = self.ctx.device
        offs = torch.empty(n + 1, dtype=torch.int32, device=dev)
        total = C.c_size_t()
        cap = 0
        while True:
            hs = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
            ts = torch.empty(max(cap, 1), dtype=torch.float32, device=dev)
            ns = torch.empty((max(cap, 1), 3), dtype=torch.float32, device=dev)
            rc = self.ctx.lib.ncx_world_interferences_with_rays(self.h, _lib.ptr(poses), _lib.ptrs(cols), n,
                                                                float(max_toi), 1 if aabb_only else 0, _lib.ptr(offs),
                                                                _lib.ptr(hs), _lib.ptr(ts), _lib.ptr(ns), cap,
                                                                C.byref(total))
            if rc == -3 and total.value > cap:
                cap = total.value
                continue
            _lib.check(rc)
            break
        k = total.value
        return offs, hs[:k], ts[:k], ns[:k]

    def interferences_with_aabbs(self, aabbs):
        """CollisionWorld::interferences_with_aabb per box of (n, 6): CSR (offsets, handles)."""
        cols = _lib.soa_columns(aabbs, 6, self.ctx.device)
        n = cols[0].numel()
        dev = self.ctx.device
        offs = torch.empty(n + 1, dtype=torch.int32, device=dev)
        total = C.c_size_t()
        cap = 0
        while True:
            hs = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
            rc = self.ctx.lib.ncx_world_interferences_with_aabbs(self.h, _lib.ptrs(cols), n, _lib.ptr(offs),
                                                                 _lib.ptr(hs), cap, C.byref(total))
            if rc == -3 and total.value > cap:
                cap = total.value
                continue
            _lib.check(rc)
            break
        return offs, hs[:total.value]

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
