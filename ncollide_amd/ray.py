"""TriMesh + RayCast (src/shape/trimesh.rs, src/query/ray/ray_trimesh.rs) on the GPU.

`TriMesh(points, indices).toi_and_normal_with_ray(...)` casts a batch of rays;
features follow the reference: face id for a front-face hit, face id + nfaces
for a back-face hit, UINT32_MAX for a miss.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from .context import default_context

NO_HIT = 0xFFFFFFFF


class TriMesh:
    def __init__(self, points, indices, ctx=None):
        self.ctx = ctx or default_context()
        self.points = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        h = C.c_void_p()
        _lib.check(self.ctx.lib.ncx_trimesh_create(self.ctx.h, _lib.ptr(self.points), len(self.points),
                                                   _lib.ptr(self.indices), len(self.indices), C.byref(h)))
        self.h = h
        self.nfaces = len(self.indices)

    def _rays(self, origins, dirs):
        dev = self.ctx.device
        o = torch.as_tensor(origins, device=dev, dtype=torch.float32).reshape(-1, 3)
        d = torch.as_tensor(dirs, device=dev, dtype=torch.float32).reshape(-1, 3)
        soa = [o[:, 0].contiguous(), o[:, 1].contiguous(), o[:, 2].contiguous(),
               d[:, 0].contiguous(), d[:, 1].contiguous(), d[:, 2].contiguous()]
        return soa

    @staticmethod
    def _pose(pose):
        p = np.zeros(8, np.float32)
        if pose is None:
            p[7] = 1.0
        else:
            p[:] = np.asarray(pose, np.float32).reshape(8)
        return p

    def toi_and_normal_with_ray(self, origins, dirs, max_toi, pose=None, soa=None):
        """Returns (toi, normal (n,3), feature) CUDA tensors.  `soa` may pass six contiguous
        device arrays (ox, oy, oz, dx, dy, dz) directly."""
        rays = soa if soa is not None else self._rays(origins, dirs)
        n = rays[0].numel()
        dev = self.ctx.device
        toi = torch.empty(n, dtype=torch.float32, device=dev)
        nr = torch.empty((n, 3), dtype=torch.float32, device=dev)
        feat = torch.empty(n, dtype=torch.int32, device=dev)
        p = self._pose(pose)
        _lib.check(self.ctx.lib.ncx_trimesh_cast_rays(self.h, _lib.ptr(p), _lib.ptr6(rays), n, float(max_toi),
                                                      _lib.ptr(toi), _lib.ptr(nr), _lib.ptr(feat)))
        return toi, nr, feat

    def cast_stats(self, origins, dirs, max_toi, pose=None, soa=None):
        """Nodes visited and triangles tested per ray by the same traversal."""
        rays = soa if soa is not None else self._rays(origins, dirs)
        n = rays[0].numel()
        dev = self.ctx.device
        nodes = torch.empty(n, dtype=torch.int32, device=dev)
        tris = torch.empty(n, dtype=torch.int32, device=dev)
        p = self._pose(pose)
        _lib.check(self.ctx.lib.ncx_trimesh_cast_stats(self.h, _lib.ptr(p), _lib.ptr6(rays), n, float(max_toi),
                                                       _lib.ptr(nodes), _lib.ptr(tris)))
        return nodes, tris

    # -- the mesh as a composite shape against convex shapes of a ShapeSet
    def contact_with_shapes(self, shapes, poses, shape_idx, prediction, pose=None):
        """contact_composite_shape_shape(this mesh at `pose`, shape shape_idx[k] at poses) per k:
        dict of CUDA tensors has_contact, world1, world2, normal (mesh toward shape), depth, face."""
        dev = self.ctx.device
        poses = torch.as_tensor(poses, device=dev, dtype=torch.float32).contiguous()
        idx = torch.as_tensor(shape_idx, device=dev).to(torch.int32).contiguous()
        n = idx.numel()
        has = torch.empty(n, dtype=torch.uint8, device=dev)
        w1, w2, nr = (torch.empty((n, 3), dtype=torch.float32, device=dev) for _ in range(3))
        depth = torch.empty(n, dtype=torch.float32, device=dev)
        face = torch.empty(n, dtype=torch.int32, device=dev)
        p = self._pose(pose)
        _lib.check(self.ctx.lib.ncx_mesh_contact_batch(self.h, _lib.ptr(p), shapes.h, _lib.ptr(poses), _lib.ptr(idx),
                                                       n, float(prediction), _lib.ptr(has), _lib.ptr(w1),
                                                       _lib.ptr(w2), _lib.ptr(nr), _lib.ptr(depth), _lib.ptr(face)))
        return dict(has_contact=has.bool(), world1=w1, world2=w2, normal=nr, depth=depth, face=face)

    def distance_to_shapes(self, shapes, poses, shape_idx, pose=None):
        """distance_composite_shape_shape(this mesh at `pose`, shape shape_idx[k]) per k (CUDA tensor)."""
        dev = self.ctx.device
        poses = torch.as_tensor(poses, device=dev, dtype=torch.float32).contiguous()
        idx = torch.as_tensor(shape_idx, device=dev).to(torch.int32).contiguous()
        n = idx.numel()
        d = torch.empty(n, dtype=torch.float32, device=dev)
        p = self._pose(pose)
        _lib.check(self.ctx.lib.ncx_mesh_distance_batch(self.h, _lib.ptr(p), shapes.h, _lib.ptr(poses),
                                                        _lib.ptr(idx), n, _lib.ptr(d)))
        return d

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.ctx.lib.ncx_trimesh_destroy(self.h)
        except Exception:
            pass
