"""Learned initial GJK direction: features, MLP training on the GPU, warm-started contacts.

The network (44 -> 64 -> 64 -> 3, FP16 operands / FP32 accumulation on the matrix cores) maps
a pair's relative pose and per-shape support samples to the `init_dir` of
contact_support_map_support_map_with_params (src/query/contact/contact_support_map_support_map.rs:38-61).
Targets are the final direction plain GJK/EPA returns for the pair (the contact normal,
expressed in shape 1's frame).  A warm-started GJK run is capped at `max_iter` iterations; a
capped pair is recomputed by plain GJK, which judges it.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib, query

NPARAMS = 8336
FEAT = 48


def descriptors(shapes):
    """(n, 16) float32 support samples of every shape of the set."""
    out = torch.empty((shapes.n, 16), dtype=torch.float32, device=shapes.ctx.device)
    _lib.check(shapes.ctx.lib.ncx_shape_descriptors(shapes.h, _lib.ptr(out)))
    return out


def _i32(x, dev):
    return torch.as_tensor(x, device=dev).to(torch.int32).contiguous()


def features(shapes, poses, desc, ia, ib):
    """(npairs, 48) float16 features."""
    dev = shapes.ctx.device
    ia, ib = _i32(ia, dev), _i32(ib, dev)
    poses = torch.as_tensor(poses, device=dev, dtype=torch.float32).contiguous()
    out = torch.empty((ia.numel(), FEAT), dtype=torch.float16, device=dev)
    _lib.check(shapes.ctx.lib.ncx_features(shapes.h, _lib.ptr(poses), _lib.ptr(desc), _lib.ptr(ia), _lib.ptr(ib),
                                           ia.numel(), _lib.ptr(out)))
    return out


class WarmStartMLP:
    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        h = C.c_void_p()
        _lib.check(ctx.lib.ncx_mlp_create(ctx.h, 44, 64, 3, C.c_uint64(seed), C.byref(h)))
        self.h = h

    def params(self):
        p = np.zeros(NPARAMS, np.float32)
        _lib.check(self.ctx.lib.ncx_mlp_get_params(self.h, _lib.ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == NPARAMS
        _lib.check(self.ctx.lib.ncx_mlp_set_params(self.h, _lib.ptr(p)))

    def forward(self, feat16):
        n = feat16.shape[0]
        out = torch.empty((n, 3), dtype=torch.float32, device=self.ctx.device)
        _lib.check(self.ctx.lib.ncx_mlp_forward(self.h, _lib.ptr(feat16), n, _lib.ptr(out)))
        return out

    def grad(self, feat16, target):
        n = feat16.shape[0]
        g = torch.empty(NPARAMS, dtype=torch.float32, device=self.ctx.device)
        loss = torch.empty(1, dtype=torch.float32, device=self.ctx.device)
        target = target.to(torch.float32).contiguous()
        _lib.check(self.ctx.lib.ncx_mlp_grad(self.h, _lib.ptr(feat16), _lib.ptr(target), n, _lib.ptr(g),
                                             _lib.ptr(loss)))
        return g, loss

    def train_step(self, feat16, target, lr, want_loss=False):
        loss = C.c_float()
        _lib.check(self.ctx.lib.ncx_mlp_train_step(self.h, _lib.ptr(feat16), _lib.ptr(target), feat16.shape[0],
                                                   float(lr), C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def predict_dirs(self, shapes, poses, desc, ia, ib):
        dev = self.ctx.device
        ia, ib = _i32(ia, dev), _i32(ib, dev)
        poses = torch.as_tensor(poses, device=dev, dtype=torch.float32).contiguous()
        out = torch.empty((ia.numel(), 3), dtype=torch.float32, device=dev)
        _lib.check(self.ctx.lib.ncx_mlp_predict_dirs(self.h, _lib.ptr(poses), _lib.ptr(desc), _lib.ptr(ia),
                                                     _lib.ptr(ib), ia.numel(), _lib.ptr(out)))
        return out

    def fit(self, feat16, target, epochs=8, batch=8192, lr=2e-3, lr_end=1e-4, seed=0, log=None):
        """Adam over shuffled mini-batches with a cosine learning-rate decay; returns per-epoch mean loss."""
        n = feat16.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        nb = (n + batch - 1) // batch
        total = epochs * nb
        hist = []
        it = 0
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(feat16.device)
            xs = feat16[perm].contiguous()
            ys = target[perm].contiguous()
            losses = []
            for b in range(nb):
                lr_t = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * it / max(total - 1, 1)))
                want = b == nb - 1 or b % 64 == 0
                l = self.train_step(xs[b * batch:(b + 1) * batch], ys[b * batch:(b + 1) * batch], lr_t,
                                    want_loss=want)
                if want:
                    losses.append(l)
                it += 1
            hist.append(float(np.mean(losses)))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4f}")
        return hist

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.ctx.lib.ncx_mlp_destroy(self.h)
        except Exception:
            pass


def frame1_targets(poses, ia, normals):
    """World contact normals -> shape-1 frame (R1^T n), float32 (n, 3) CUDA tensor."""
    q = poses[ia.long(), 4:8]
    qv, w = -q[:, :3], q[:, 3:4]
    t = 2.0 * torch.cross(qv, normals, dim=1)
    return (normals + w * t + torch.cross(qv, t, dim=1)).contiguous()


def warm_contacts(mlp, shapes, poses, desc, ia, ib, prediction=0.0, max_iter=32):
    """query::contact with the learned init_dir (capped GJK, plain GJK on the cap)."""
    dirs = mlp.predict_dirs(shapes, poses, desc, ia, ib)
    return query.contact(shapes, poses, ia, ib, prediction, init_dir=dirs, max_iter=max_iter), dirs
