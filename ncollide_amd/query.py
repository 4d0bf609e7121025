"""Batched query::contact / query::distance (src/query/contact/contact_shape_shape.rs:10-48,
src/query/distance/distance.rs:9-33) over pairs of a ShapeSet.

Each pair (ia[k], ib[k]) is evaluated exactly as the reference's per-pair call
`query::contact(&poses[ia], shape[ia], &poses[ib], shape[ib], prediction)`.
Inputs and outputs are CUDA tensors; poses are (n_shapes, 8) float32 in the
ncx_pose layout.
"""
import torch

from . import _lib


def _u32(t, dev):
    return torch.as_tensor(t, device=dev).to(torch.int32).contiguous()


def contact(shapes, poses, ia, ib, prediction, init_dir=None, max_iter=None):
    """Returns a dict of CUDA tensors: has_contact (bool), world1, world2, normal (n,3), depth, gjk_iters.

    init_dir (n,3), optional: contact_support_map_support_map_with_params's
    `init_dir` for pairs on the support-map path (None = the reference's None).
    max_iter, optional: GJK iteration cap for the init_dir run; a capped pair is
    recomputed by plain GJK.
    """
    ctx = shapes.ctx
    dev = ctx.device
    poses = torch.as_tensor(poses, device=dev, dtype=torch.float32).contiguous()
    ia = _u32(ia, dev)
    ib = _u32(ib, dev)
    n = ia.numel()
    if ib.numel() != n:
        raise ValueError("ia and ib differ in length")
    has = torch.empty(n, dtype=torch.uint8, device=dev)
    w1 = torch.empty((n, 3), dtype=torch.float32, device=dev)
    w2 = torch.empty((n, 3), dtype=torch.float32, device=dev)
    nr = torch.empty((n, 3), dtype=torch.float32, device=dev)
    depth = torch.empty(n, dtype=torch.float32, device=dev)
    iters = torch.empty(n, dtype=torch.int32, device=dev)
    idir = None
    if init_dir is not None:
        idir = torch.as_tensor(init_dir, device=dev, dtype=torch.float32).contiguous()
    if max_iter is None:
        _lib.check(ctx.lib.ncx_contact_batch(shapes.h, _lib.ptr(poses), _lib.ptr(ia), _lib.ptr(ib), n,
                                             float(prediction), _lib.ptr(idir), _lib.ptr(has), _lib.ptr(w1),
                                             _lib.ptr(w2), _lib.ptr(nr), _lib.ptr(depth), _lib.ptr(iters)))
    else:
        _lib.check(ctx.lib.ncx_contact_batch_capped(shapes.h, _lib.ptr(poses), _lib.ptr(ia), _lib.ptr(ib), n,
                                                    float(prediction), _lib.ptr(idir), int(max_iter), _lib.ptr(has),
                                                    _lib.ptr(w1), _lib.ptr(w2), _lib.ptr(nr), _lib.ptr(depth),
                                                    _lib.ptr(iters)))
    return dict(has_contact=has.bool(), world1=w1, world2=w2, normal=nr, depth=depth, gjk_iters=iters)


def distance(shapes, poses, ia, ib, init_dir=None, max_iter=None):
    """Minimum distance per pair (0 when touching or penetrating) and GJK iterations.

    init_dir (n,3), optional: distance_support_map_support_map_with_params's `init_dir`
    (None, or an all-zero row, = the reference's None).  max_iter, optional: GJK iteration
    cap for the init_dir run; a capped pair is recomputed by plain GJK.
    """
    ctx = shapes.ctx
    dev = ctx.device
    poses = torch.as_tensor(poses, device=dev, dtype=torch.float32).contiguous()
    ia = _u32(ia, dev)
    ib = _u32(ib, dev)
    n = ia.numel()
    if ib.numel() != n:
        raise ValueError("ia and ib differ in length")
    d = torch.empty(n, dtype=torch.float32, device=dev)
    iters = torch.empty(n, dtype=torch.int32, device=dev)
    if init_dir is None and max_iter is None:
        _lib.check(ctx.lib.ncx_distance_batch(shapes.h, _lib.ptr(poses), _lib.ptr(ia), _lib.ptr(ib), n,
                                              _lib.ptr(d), _lib.ptr(iters)))
    else:
        idir = None
        if init_dir is not None:
            idir = torch.as_tensor(init_dir, device=dev, dtype=torch.float32).contiguous()
        _lib.check(ctx.lib.ncx_distance_batch_capped(shapes.h, _lib.ptr(poses), _lib.ptr(ia), _lib.ptr(ib), n,
                                                     _lib.ptr(idir), int(10000 if max_iter is None else max_iter),
                                                     _lib.ptr(d), _lib.ptr(iters)))
    return d, iters


def time_of_impact(shapes, poses, velocities, ia, ib, max_toi, target_distance=0.0):
    """Batched query::time_of_impact (src/query/time_of_impact/time_of_impact.rs:66-158) under
    translational motion: velocities (n_shapes, 3) per shape, like poses.

    Returns a dict of CUDA tensors: has (bool), toi, witness1, witness2, normal1, normal2 (n,3,
    local frames of each pair's shapes) and status (0 OutOfIterations, 1 Converged, 2 Failed,
    3 Penetrating).
    """
    ctx = shapes.ctx
    dev = ctx.device
    poses = torch.as_tensor(poses, device=dev, dtype=torch.float32).contiguous()
    vel = torch.as_tensor(velocities, device=dev, dtype=torch.float32).reshape(-1, 3).contiguous()
    if vel.shape[0] != poses.shape[0]:
        raise ValueError("one velocity per shape is required")
    ia = _u32(ia, dev)
    ib = _u32(ib, dev)
    n = ia.numel()
    if ib.numel() != n:
        raise ValueError("ia and ib differ in length")
    has = torch.empty(n, dtype=torch.uint8, device=dev)
    toi = torch.empty(n, dtype=torch.float32, device=dev)
    w1, w2, n1, n2 = (torch.empty((n, 3), dtype=torch.float32, device=dev) for _ in range(4))
    status = torch.empty(n, dtype=torch.uint8, device=dev)
    _lib.check(ctx.lib.ncx_time_of_impact_batch(shapes.h, _lib.ptr(poses), _lib.ptr(vel), _lib.ptr(ia), _lib.ptr(ib),
                                                n, float(max_toi), float(target_distance), _lib.ptr(has),
                                                _lib.ptr(toi), _lib.ptr(w1), _lib.ptr(w2), _lib.ptr(n1), _lib.ptr(n2),
                                                _lib.ptr(status)))
    return dict(has=has.bool(), toi=toi, witness1=w1, witness2=w2, normal1=n1, normal2=n2, status=status)


def cast_rays(shapes, poses, shape_idx, origins, dirs, max_toi, solid=True):
    """Batched RayCast::toi_and_normal_with_ray (src/query/ray/ray.rs:125-160) of shape
    shape_idx[k] at its pose against ray k.  Returns (toi, normal (n,3), feature) CUDA tensors;
    a miss has toi -1 and feature 0xFFFFFFFF (as int32: -1); FeatureId::Unknown is 0xFFFFFFFE."""
    ctx = shapes.ctx
    dev = ctx.device
    poses = torch.as_tensor(poses, device=dev, dtype=torch.float32).contiguous()
    idx = _u32(shape_idx, dev)
    o = torch.as_tensor(origins, device=dev, dtype=torch.float32).reshape(-1, 3)
    d = torch.as_tensor(dirs, device=dev, dtype=torch.float32).reshape(-1, 3)
    n = idx.numel()
    if o.shape[0] != n or d.shape[0] != n:
        raise ValueError("one ray per shape index is required")
    soa = [o[:, 0].contiguous(), o[:, 1].contiguous(), o[:, 2].contiguous(),
           d[:, 0].contiguous(), d[:, 1].contiguous(), d[:, 2].contiguous()]
    toi = torch.empty(n, dtype=torch.float32, device=dev)
    nrm = torch.empty((n, 3), dtype=torch.float32, device=dev)
    feat = torch.empty(n, dtype=torch.int32, device=dev)
    _lib.check(ctx.lib.ncx_shapes_cast_rays(shapes.h, _lib.ptr(poses), _lib.ptr(idx), _lib.ptr6(soa), n,
                                            float(max_toi), int(bool(solid)), _lib.ptr(toi), _lib.ptr(nrm),
                                            _lib.ptr(feat)))
    return toi, nrm, feat
