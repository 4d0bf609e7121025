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
