"""Persistent contact manifolds: the narrow phase's contact-manifold generators for Ball, Cuboid,
Capsule and ConvexHull pairs, batched (ncx_contact_manifolds).

Mirrors NarrowPhase::update_contact (src/pipeline/narrow_phase/narrow_phase.rs:56-80) with the
generators DefaultContactDispatcher picks (default_contact_dispatcher.rs:27-97) and the
ContactManifold they fill (src/query/contact/contact_manifold.rs): one slot per interaction keeps
the polyhedron generator's last_gjk_dir and the manifold's tracked contact points across updates.
"""
import torch

from . import _lib

MAX_CONTACTS = 32  # NCX_MANIFOLD_MAX

# FeatureId encoding: kind in the top two bits, the shape-local id below.
VERTEX, EDGE, FACE, UNKNOWN = 0, 1, 2, 3


def feature_kind(f):
    return (f >> 30) & 3


def feature_id(f):
    return f & 0x3FFFFFFF


class ContactManifolds:
    """State of `nslots` interactions (a generator and its ContactManifold each)."""

    def __init__(self, ctx, nslots):
        import ctypes as C
        self.ctx = ctx
        self.nslots = int(nslots)
        h = C.c_void_p()
        _lib.check(ctx.lib.ncx_manifolds_create(ctx.h, self.nslots, C.byref(h)))
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.ctx.lib.ncx_manifolds_destroy(h)
            self.h = None

    def _slots(self, slots):
        return torch.as_tensor(slots, device=self.ctx.device).to(torch.int32).contiguous()

    def reset(self, slots):
        """New interactions: forget last_gjk_dir and the tracked contacts of these slots."""
        s = self._slots(slots)
        _lib.check(self.ctx.lib.ncx_manifolds_reset(self.h, _lib.ptr(s), s.numel()))

    def dirs(self, slots):
        """last_gjk_dir per slot, (n, 3) CUDA tensor; zero rows: None."""
        s = self._slots(slots)
        out = torch.zeros((s.numel(), 3), dtype=torch.float32, device=self.ctx.device)
        _lib.check(self.ctx.lib.ncx_manifolds_dirs(self.h, _lib.ptr(s), s.numel(), _lib.ptr(out)))
        return out

    def update(self, shapes, poses, ia, ib, slots, linear, angular1=0.0, angular2=0.0, out=None):
        """One generate_contacts per pair (ia[k], ib[k]) on slot slots[k] (distinct within a
        call).  Returns CUDA tensors: count (n,), and per pair MAX_CONTACTS rows of world1,
        world2, normal (n, MAX, 3), depth, feature1, feature2 (n, MAX); rows past count[k] are
        unused.  Contacts are in the manifold's cache order.  out: a previous result for the same
        number of pairs, whose tensors are written again."""
        dev = self.ctx.device
        poses = torch.as_tensor(poses, device=dev, dtype=torch.float32).contiguous()
        ia = torch.as_tensor(ia, device=dev).to(torch.int32).contiguous()
        ib = torch.as_tensor(ib, device=dev).to(torch.int32).contiguous()
        s = self._slots(slots)
        n = ia.numel()
        if ib.numel() != n or s.numel() != n:
            raise ValueError("ia, ib and slots differ in length")
        m = MAX_CONTACTS
        if out is not None:
            if out["count"].numel() != n:
                raise ValueError("out holds another number of pairs")
            count, w1, w2, nr = out["count"], out["world1"], out["world2"], out["normal"]
            depth, f1, f2 = out["depth"], out["feature1"], out["feature2"]
        else:
            count = torch.zeros(n, dtype=torch.int32, device=dev)
            w1 = torch.zeros((n, m, 3), dtype=torch.float32, device=dev)
            w2 = torch.zeros((n, m, 3), dtype=torch.float32, device=dev)
            nr = torch.zeros((n, m, 3), dtype=torch.float32, device=dev)
            depth = torch.zeros((n, m), dtype=torch.float32, device=dev)
            f1 = torch.zeros((n, m), dtype=torch.int32, device=dev)
            f2 = torch.zeros((n, m), dtype=torch.int32, device=dev)
        p = _lib.ptr
        _lib.check(self.ctx.lib.ncx_contact_manifolds(self.h, shapes.h, p(poses), p(ia), p(ib), p(s), n,
                                                      float(linear), float(angular1), float(angular2), p(count),
                                                      p(w1), p(w2), p(nr), p(depth), p(f1), p(f2)))
        return dict(count=count, world1=w1, world2=w2, normal=nr, depth=depth, feature1=f1, feature2=f2)
