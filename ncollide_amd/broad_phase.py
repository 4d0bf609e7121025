"""DBVTBroadPhase (src/pipeline/broad_phase/dbvt_broad_phase.rs) on the GPU.

Same method names and meaning as the reference's BroadPhase trait
(src/pipeline/broad_phase/broad_phase.rs:41-86); boxes are 6 floats
(minx, miny, minz, maxx, maxy, maxz).  Batch variants take (n, 6) arrays or
six device arrays.  Pair keys are (a << 32) | b with a < b.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from .context import default_context


class DBVTBroadPhase:
    def __init__(self, margin, ctx=None):
        self.ctx = ctx or default_context()
        if margin < 0:
            raise ValueError("The loosening margin must be positive.")
        h = C.c_void_p()
        _lib.check(self.ctx.lib.ncx_bp_create(self.ctx.h, float(margin), C.byref(h)))
        self.h = h
        self.margin = float(margin)

    # -- proxies
    def create_proxy(self, aabb):
        return int(self.create_proxies(np.asarray(aabb, np.float32).reshape(1, 6))[0])

    def create_proxies(self, aabbs):
        a = np.ascontiguousarray(np.asarray(aabbs, np.float32).reshape(-1, 6).T)
        n = a.shape[1]
        out = np.zeros(n, np.uint32)
        cols = [np.ascontiguousarray(a[d]) for d in range(6)]
        _lib.check(self.ctx.lib.ncx_bp_create_proxies(self.h, _lib.ptr6(cols), n, _lib.ptr(out)))
        return out

    def deferred_set_bounding_volume(self, handle, aabb):
        self.deferred_set_bounding_volumes(np.array([handle], np.uint32), np.asarray(aabb, np.float32).reshape(1, 6))

    def deferred_set_bounding_volumes(self, handles, aabbs):
        """Host arrays: handles (n,), boxes (n, 6).  Handles within one call must be distinct."""
        hs = np.ascontiguousarray(handles, np.uint32)
        a = np.ascontiguousarray(np.asarray(aabbs, np.float32).reshape(-1, 6).T)
        cols = [np.ascontiguousarray(a[d]) for d in range(6)]
        _lib.check(self.ctx.lib.ncx_bp_set_aabbs(self.h, _lib.ptr(hs), _lib.ptr6(cols), len(hs)))

    def deferred_set_bounding_volumes_dev(self, handles_dev, aabb_soa_dev):
        """Device variant: handles (n,) int32/uint32 tensor, aabb_soa_dev = six (n,) float32 tensors."""
        _lib.check(self.ctx.lib.ncx_bp_set_aabbs_dev(self.h, _lib.ptr(handles_dev), _lib.ptr6(aabb_soa_dev),
                                                     handles_dev.numel()))

    def remove(self, handles):
        """BroadPhase::remove: returns the sorted keys of the pairs it drops (the reference hands
        each to its handler); they are not reported again as stopped by the next update."""
        hs = np.ascontiguousarray(handles, np.uint32)
        n = C.c_size_t()
        cap = max(1, self.num_interferences())
        out = np.zeros(cap, np.uint64)
        _lib.check(self.ctx.lib.ncx_bp_remove_with_pairs(self.h, _lib.ptr(hs), len(hs), _lib.ptr(out), cap,
                                                         C.byref(n)))
        return out[:n.value]

    def set_collision_groups(self, handles, membership, whitelist, blacklist):
        """Per-proxy CollisionGroups masks, the pair filter (can_interact_with_groups)."""
        hs = np.ascontiguousarray(handles, np.uint32)
        m = np.ascontiguousarray(membership, np.uint32)
        w = np.ascontiguousarray(whitelist, np.uint32)
        b = np.ascontiguousarray(blacklist, np.uint32)
        _lib.check(self.ctx.lib.ncx_bp_set_collision_groups(self.h, _lib.ptr(hs), _lib.ptr(m), _lib.ptr(w),
                                                            _lib.ptr(b), len(hs)))

    # -- interference queries (batched; CSR results: offsets (nq + 1,), handles)
    def _collect(self, fn, cols, nq, *extra):
        dev = self.ctx.device
        offs = torch.empty(nq + 1, dtype=torch.int32, device=dev)
        total = C.c_size_t()
        cap = 0
        handles = torch.empty(1, dtype=torch.int32, device=dev)
        while True:
            rc = fn(self.h, _lib.ptrs(cols), nq, *extra, _lib.ptr(offs), _lib.ptr(handles), cap, C.byref(total))
            if rc == -3 and total.value > cap:
                cap = total.value
                handles = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
                continue
            _lib.check(rc)
            break
        return offs, handles[:total.value]

    def interferences_with_bounding_volumes(self, aabbs):
        """interferences_with_bounding_volume for each box of (nq, 6): (offsets, handles) CUDA tensors."""
        cols = _lib.soa_columns(aabbs, 6, self.ctx.device)
        return self._collect(self.ctx.lib.ncx_bp_interferences_with_aabbs, cols, cols[0].numel())

    def interferences_with_points(self, points):
        cols = _lib.soa_columns(points, 3, self.ctx.device)
        return self._collect(self.ctx.lib.ncx_bp_interferences_with_points, cols, cols[0].numel())

    def interferences_with_rays(self, origins, dirs, max_toi):
        cols = _lib.soa_columns(np.hstack([np.asarray(origins, np.float32).reshape(-1, 3),
                                           np.asarray(dirs, np.float32).reshape(-1, 3)]), 6, self.ctx.device)
        return self._collect(self.ctx.lib.ncx_bp_interferences_with_rays, cols, cols[0].numel(), C.c_float(max_toi))

    # -- update and results
    def update(self):
        _lib.check(self.ctx.lib.ncx_bp_update(self.h))

    def num_interferences(self):
        n = C.c_size_t()
        _lib.check(self.ctx.lib.ncx_bp_num_pairs(self.h, C.byref(n)))
        return n.value

    def pairs(self):
        """Sorted pair keys (numpy uint64)."""
        n = self.num_interferences()
        out = np.zeros(max(n, 1), np.uint64)
        got = C.c_size_t()
        _lib.check(self.ctx.lib.ncx_bp_pairs(self.h, _lib.ptr(out), len(out), C.byref(got)))
        return out[:got.value]

    def pairs_tensors(self):
        """The current pair list (unsorted, a < b) as two int32 CUDA tensors."""
        n = self.num_interferences()
        a = torch.empty(max(n, 1), dtype=torch.int32, device=self.ctx.device)
        b = torch.empty(max(n, 1), dtype=torch.int32, device=self.ctx.device)
        got = C.c_size_t()
        _lib.check(self.ctx.lib.ncx_bp_copy_pairs_dev(self.h, _lib.ptr(a), _lib.ptr(b), n, C.byref(got)))
        return a[:n], b[:n]

    def events(self):
        """(started, stopped) sorted pair keys of the last update."""
        ns = C.c_size_t()
        nt = C.c_size_t()
        rc = self.ctx.lib.ncx_bp_events(self.h, None, 0, C.byref(ns), None, 0, C.byref(nt))
        if rc not in (0, -3):
            _lib.check(rc)
        s = np.zeros(max(ns.value, 1), np.uint64)
        t = np.zeros(max(nt.value, 1), np.uint64)
        _lib.check(self.ctx.lib.ncx_bp_events(self.h, _lib.ptr(s), len(s), C.byref(ns), _lib.ptr(t), len(t),
                                              C.byref(nt)))
        return s[:ns.value], t[:nt.value]

    def proxy(self, handle):
        out = np.zeros(6, np.float32)
        _lib.check(self.ctx.lib.ncx_bp_proxy_aabb(self.h, int(handle), _lib.ptr(out)))
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.ctx.lib.ncx_bp_destroy(self.h)
        except Exception:
            pass


def pair_keys_to_arrays(keys):
    keys = np.asarray(keys, np.uint64)
    return (keys >> np.uint64(32)).astype(np.uint32), (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def pair_keys_from_tensors(a, b):
    a = a.to(torch.int64).cpu().numpy().astype(np.uint64)
    b = b.to(torch.int64).cpu().numpy().astype(np.uint64)
    return np.sort((a << np.uint64(32)) | b)
