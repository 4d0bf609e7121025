"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The product package
(ncollide_amd) never imports it.

Shape sets use the same layout as the product: types (0 ball, 1 cuboid,
2 capsule, 3 hull), params (3 per shape), hull_off/hull_cnt into a hull point
pool (xyz per point).  Poses are (tx, ty, tz, qi, qj, qk, qw).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
I64 = C.c_int64


def _declare(L):
    def sig(name, res, *args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = list(args)

    for t, R in (("f32", C.c_float), ("f64", C.c_double)):
        sig(f"orc_contact_{t}", C.c_int, I64, P, P, P, P, P, I64, P, P, P, I64, R, P, P, P, P, P, P, P, C.c_int)
        sig(f"orc_contact_sm_{t}", C.c_int, I64, P, P, P, P, P, I64, P, P, P, I64, R, P, C.c_int32,
            P, P, P, P, P, P, P, P, C.c_int)
        sig(f"orc_distance_{t}", C.c_int, I64, P, P, P, P, P, I64, P, P, P, I64, P, C.c_int32, P, P, C.c_int)
        sig(f"orc_shape_aabb_{t}", C.c_int, I64, P, P, P, P, P, I64, P, P)
        sig(f"orc_trimesh_create_{t}", P, P, I64, P, I64)
        sig(f"orc_trimesh_free_{t}", None, P)
        sig(f"orc_trimesh_cast_{t}", C.c_int, P, P, P, P, I64, R, P, P, P, P, C.c_int)
        sig(f"orc_trimesh_brute_{t}", C.c_int, P, P, I64, P, P, P, I64, R, P, P, P, P, C.c_int)
        sig(f"orc_toi_{t}", C.c_int, I64, P, P, P, P, P, I64, P, P, P, P, I64, R, R, P, P, P, P, P, P, P, P,
            C.c_int)
        sig(f"orc_shape_rays_{t}", C.c_int, I64, P, P, P, P, P, I64, P, P, P, P, I64, R, C.c_int, P, P, P, P, C.c_int)
    sig("orc_world_rays_f32", I64, I64, P, P, P, P, P, I64, P, P, P, P, I64, C.c_float, C.c_int, P, P, P, P, I64,
        C.c_int)
    sig("orc_manifold_store_new", P, I64)
    sig("orc_manifold_store_free", None, P)
    sig("orc_manifold_reset", C.c_int, P, P, I64)
    sig("orc_manifold_step_f32", C.c_int, P, I64, P, P, P, P, P, I64, P, P, P, P, P, P, P, I64, C.c_float,
        C.c_float, C.c_float, I64, P, P, P, P, P, P, P, P, C.c_int)
    sig("orc_hull_polyhedron_f32", C.c_int, P, I64, P, I64, P, P, P, P, P, P, P, P, P, P, P, P, P, P)
    sig("orc_polyhedra_step_f32", C.c_int, I64, P, P, P, P, P, I64, P, P, P, I64, C.c_float, P, P, P, P, P, P, P,
        P, C.c_int)
    for t, R in (("f32", C.c_float), ("f64", C.c_double)):
        sig(f"orc_mesh_queries_{t}", C.c_int, P, I64, P, I64, P, P, P, P, P, I64, P, P, I64, R, C.c_int, P, P, P, P, P,
            P, P, C.c_int)
    sig("orc_ray_triangle_f64", C.c_int, P, P, P, P, P, P)
    sig("orc_ray_aabb_f64", C.c_int, P, P, P, C.c_double, C.c_int, P)
    sig("orc_project_f64", C.c_int, C.c_int, P, P, P, P, P)
    sig("orc_orthonormal_first_f64", None, P, P)
    sig("orc_affinely_dependent_f64", C.c_int, P)
    sig("orc_affinely_dependent_f32", C.c_int, P)
    sig("orc_bp_create", P, C.c_double)
    sig("orc_bp_free", None, P)
    sig("orc_bp_create_proxy", C.c_uint32, P, P)
    sig("orc_bp_set_bv", None, P, C.c_uint32, P)
    sig("orc_bp_remove", None, P, P, I64)
    sig("orc_bp_update", None, P)
    sig("orc_bp_set_groups", None, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32)
    sig("orc_bp_num_pairs", I64, P)
    sig("orc_bp_pairs", I64, P, P, I64)
    sig("orc_bp_events", I64, P, C.c_int, P, I64)
    sig("orc_bp_clear_events", None, P)
    sig("orc_bp_stored_bv", C.c_int, P, C.c_uint32, P)
    sig("orc_pairs_brute", I64, P, I64, P, I64)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _dt(dtype):
    return ("f32", np.float32) if np.dtype(dtype) == np.float32 else ("f64", np.float64)


class ShapeSet:
    """Host arrays describing shapes, in the oracle's scalar type."""

    def __init__(self, types, params, hull_off, hull_cnt, hull_pts, dtype=np.float32):
        t, R = _dt(dtype)
        self.tag = t
        self.R = R
        self.types = np.ascontiguousarray(types, dtype=np.uint32)
        self.params = np.ascontiguousarray(params, dtype=R).reshape(-1, 3)
        self.hull_off = np.ascontiguousarray(hull_off, dtype=np.uint32)
        self.hull_cnt = np.ascontiguousarray(hull_cnt, dtype=np.uint32)
        hp = np.asarray(hull_pts, dtype=R).reshape(-1, 3)
        if hp.shape[0] == 0:
            hp = np.zeros((1, 3), dtype=R)
        self.hull_pts = np.ascontiguousarray(hp)
        self.n = len(self.types)

    def args(self):
        return (self.n, _p(self.types), _p(self.params), _p(self.hull_off), _p(self.hull_cnt),
                _p(self.hull_pts), self.hull_pts.shape[0])


def contact(ss, poses, ia, ib, prediction, nthreads=1):
    L = lib()
    R = ss.R
    poses = np.ascontiguousarray(poses, dtype=R).reshape(-1, 7)
    ia = np.ascontiguousarray(ia, dtype=np.uint32)
    ib = np.ascontiguousarray(ib, dtype=np.uint32)
    n = len(ia)
    has = np.zeros(n, np.uint8)
    w1 = np.zeros((n, 3), R); w2 = np.zeros((n, 3), R); nr = np.zeros((n, 3), R)
    depth = np.zeros(n, R)
    gi = np.zeros(n, np.int32); ei = np.zeros(n, np.int32)
    getattr(L, f"orc_contact_{ss.tag}")(*ss.args(), _p(poses), _p(ia), _p(ib), n, prediction, _p(has), _p(w1),
                                         _p(w2), _p(nr), _p(depth), _p(gi), _p(ei), nthreads)
    return dict(has=has.astype(bool), world1=w1, world2=w2, normal=nr, depth=depth, gjk_iters=gi, epa_iters=ei)


def contact_support_map(ss, poses, ia, ib, prediction, init_dirs=None, max_iter=10000, nthreads=1):
    L = lib()
    R = ss.R
    poses = np.ascontiguousarray(poses, dtype=R).reshape(-1, 7)
    ia = np.ascontiguousarray(ia, dtype=np.uint32)
    ib = np.ascontiguousarray(ib, dtype=np.uint32)
    n = len(ia)
    dirs = None if init_dirs is None else np.ascontiguousarray(init_dirs, dtype=R).reshape(-1, 3)
    has = np.zeros(n, np.uint8)
    w1 = np.zeros((n, 3), R); w2 = np.zeros((n, 3), R); nr = np.zeros((n, 3), R)
    depth = np.zeros(n, R)
    gi = np.zeros(n, np.int32); ei = np.zeros(n, np.int32); st = np.zeros(n, np.int32)
    getattr(L, f"orc_contact_sm_{ss.tag}")(*ss.args(), _p(poses), _p(ia), _p(ib), n, prediction, _p(dirs),
                                            max_iter, _p(has), _p(w1), _p(w2), _p(nr), _p(depth), _p(gi), _p(ei),
                                            _p(st), nthreads)
    return dict(has=has.astype(bool), world1=w1, world2=w2, normal=nr, depth=depth, gjk_iters=gi, epa_iters=ei,
                status=st)


def distance(ss, poses, ia, ib, nthreads=1, init_dirs=None, max_iter=10000):
    """query::distance per pair; init_dirs (n,3) optional (all-zero row = None), GJK from
    init_dir capped at max_iter with plain GJK on the cap."""
    L = lib()
    R = ss.R
    poses = np.ascontiguousarray(poses, dtype=R).reshape(-1, 7)
    ia = np.ascontiguousarray(ia, dtype=np.uint32)
    ib = np.ascontiguousarray(ib, dtype=np.uint32)
    n = len(ia)
    d = np.zeros(n, R)
    gi = np.zeros(n, np.int32)
    dirs = None if init_dirs is None else np.ascontiguousarray(init_dirs, dtype=R).reshape(n, 3)
    getattr(L, f"orc_distance_{ss.tag}")(*ss.args(), _p(poses), _p(ia), _p(ib), n, _p(dirs), int(max_iter), _p(d),
                                         _p(gi), nthreads)
    return d, gi


def shape_aabbs(ss, poses):
    L = lib()
    poses = np.ascontiguousarray(poses, dtype=ss.R).reshape(-1, 7)
    out = np.zeros((ss.n, 6), ss.R)
    getattr(L, f"orc_shape_aabb_{ss.tag}")(*ss.args(), _p(poses), _p(out))
    return out


class TriMesh:
    """The reference TriMesh: median-split BVT + best-first ray cast."""

    def __init__(self, points, indices, dtype=np.float32):
        self.tag, self.R = _dt(dtype)
        self.points = np.ascontiguousarray(points, dtype=self.R).reshape(-1, 3)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        self.h = getattr(lib(), f"orc_trimesh_create_{self.tag}")(_p(self.points), self.points.shape[0],
                                                                    _p(self.indices), self.indices.shape[0])

    def __del__(self):
        try:
            getattr(lib(), f"orc_trimesh_free_{self.tag}")(self.h)
        except Exception:
            pass

    def cast(self, origins, dirs, max_toi, pose=None, nthreads=1):
        R = self.R
        pose = np.array([0, 0, 0, 0, 0, 0, 1], R) if pose is None else np.ascontiguousarray(pose, dtype=R)
        o = np.ascontiguousarray(origins, dtype=R).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=R).reshape(-1, 3)
        n = o.shape[0]
        toi = np.zeros(n, R); nr = np.zeros((n, 3), R)
        feat = np.zeros(n, np.uint32); visits = np.zeros(n, np.uint32)
        getattr(lib(), f"orc_trimesh_cast_{self.tag}")(self.h, _p(pose), _p(o), _p(d), n, max_toi, _p(toi),
                                                       _p(nr), _p(feat), _p(visits), nthreads)
        return toi, nr, feat, visits

    def brute(self, origins, dirs, max_toi, pose=None, nthreads=1):
        R = self.R
        pose = np.array([0, 0, 0, 0, 0, 0, 1], R) if pose is None else np.ascontiguousarray(pose, dtype=R)
        o = np.ascontiguousarray(origins, dtype=R).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=R).reshape(-1, 3)
        n = o.shape[0]
        toi = np.zeros(n, R); nr = np.zeros((n, 3), R)
        feat = np.zeros(n, np.uint32); nh = np.zeros(n, np.uint32)
        getattr(lib(), f"orc_trimesh_brute_{self.tag}")(_p(self.points), _p(self.indices), self.indices.shape[0],
                                                        _p(pose), _p(o), _p(d), n, max_toi, _p(toi), _p(nr),
                                                        _p(feat), _p(nh), nthreads)
        return toi, nr, feat, nh


def ray_triangle(abc, origin, direction):
    abc = np.ascontiguousarray(abc, np.float64).reshape(9)
    o = np.ascontiguousarray(origin, np.float64); d = np.ascontiguousarray(direction, np.float64)
    toi = np.zeros(1); n = np.zeros(3); fid = np.zeros(1, np.int32)
    ok = lib().orc_ray_triangle_f64(_p(abc), _p(o), _p(d), _p(toi), _p(n), _p(fid))
    return (bool(ok), toi[0], n, int(fid[0]))


def ray_aabb(bb6, origin, direction, max_toi, solid=True):
    bb = np.ascontiguousarray(bb6, np.float64)
    o = np.ascontiguousarray(origin, np.float64); d = np.ascontiguousarray(direction, np.float64)
    toi = np.zeros(1)
    ok = lib().orc_ray_aabb_f64(_p(bb), _p(o), _p(d), max_toi, 1 if solid else 0, _p(toi))
    return bool(ok), toi[0]


def project_origin(points):
    """Voronoi-region projection of the origin on a segment/triangle/tetrahedron (f64)."""
    pts = np.zeros((4, 3))
    pts[: len(points)] = points
    pts = np.ascontiguousarray(pts)
    proj = np.zeros(3); lk = np.zeros(1, np.int32); li = np.zeros(1, np.int32); bc = np.zeros(3)
    lib().orc_project_f64(len(points) - 1, _p(pts), _p(proj), _p(lk), _p(li), _p(bc))
    return proj, int(lk[0]), int(li[0]), bc


def orthonormal_subspace_first(v):
    """The direction EPA's 1-simplex start adds its third point along (f64)."""
    v = np.ascontiguousarray(v, np.float64)
    out = np.zeros(3)
    lib().orc_orthonormal_first_f64(_p(v), _p(out))
    return out


def is_affinely_dependent_triangle(p1, p2, p3, dtype=np.float64):
    pts = np.ascontiguousarray(np.concatenate([p1, p2, p3]), dtype)
    f = lib().orc_affinely_dependent_f32 if np.dtype(dtype) == np.float32 else lib().orc_affinely_dependent_f64
    return bool(f(_p(pts)))


class BroadPhase:
    """The reference DBVTBroadPhase restated (f32 AABBs)."""

    def __init__(self, margin):
        self.h = lib().orc_bp_create(margin)

    def __del__(self):
        try:
            lib().orc_bp_free(self.h)
        except Exception:
            pass

    def create_proxy(self, bb6):
        b = np.ascontiguousarray(bb6, np.float32)
        return int(lib().orc_bp_create_proxy(self.h, _p(b)))

    def deferred_set_bounding_volume(self, handle, bb6):
        b = np.ascontiguousarray(bb6, np.float32)
        lib().orc_bp_set_bv(self.h, handle, _p(b))

    def remove(self, handles):
        hs = np.ascontiguousarray(handles, np.uint32)
        lib().orc_bp_remove(self.h, _p(hs), len(hs))

    def update(self):
        lib().orc_bp_update(self.h)

    def set_collision_groups(self, handle, membership, whitelist, blacklist):
        lib().orc_bp_set_groups(self.h, int(handle), int(membership), int(whitelist), int(blacklist))

    def num_interferences(self):
        return int(lib().orc_bp_num_pairs(self.h))

    def pairs(self):
        n = self.num_interferences()
        out = np.zeros(max(n, 1), np.uint64)
        lib().orc_bp_pairs(self.h, _p(out), n)
        return out[:n]

    def events(self):
        res = []
        for which in (0, 1):
            n = lib().orc_bp_events(self.h, which, None, 0)
            out = np.zeros(max(n, 1), np.uint64)
            lib().orc_bp_events(self.h, which, _p(out), n)
            res.append(out[:n])
        lib().orc_bp_clear_events(self.h)
        return res[0], res[1]

    def stored_bv(self, handle):
        out = np.zeros(6, np.float32)
        ok = lib().orc_bp_stored_bv(self.h, handle, _p(out))
        return out if ok else None


def pairs_brute(bb6):
    bb = np.ascontiguousarray(bb6, np.float32).reshape(-1, 6)
    n = lib().orc_pairs_brute(_p(bb), bb.shape[0], None, 0)
    out = np.zeros(max(n, 1), np.uint64)
    lib().orc_pairs_brute(_p(bb), bb.shape[0], _p(out), n)
    return out[:n]


def time_of_impact(ss, poses, vel, ia, ib, max_toi, target_distance, nthreads=1):
    """query::time_of_impact per pair; velocities per shape (n, 3)."""
    L = lib()
    R = ss.R
    poses = np.ascontiguousarray(poses, dtype=R).reshape(-1, 7)
    vel = np.ascontiguousarray(vel, dtype=R).reshape(-1, 3)
    ia = np.ascontiguousarray(ia, dtype=np.uint32)
    ib = np.ascontiguousarray(ib, dtype=np.uint32)
    n = len(ia)
    has = np.zeros(n, np.uint8)
    toi = np.zeros(n, R)
    w1, w2, n1, n2 = (np.zeros((n, 3), R) for _ in range(4))
    status = np.zeros(n, np.uint8)
    it = np.zeros(n, np.int32)
    getattr(L, f"orc_toi_{ss.tag}")(*ss.args(), _p(poses), _p(vel), _p(ia), _p(ib), n, max_toi, target_distance,
                                     _p(has), _p(toi), _p(w1), _p(w2), _p(n1), _p(n2), _p(status), _p(it), nthreads)
    return dict(has=has.astype(bool), toi=toi, witness1=w1, witness2=w2, normal1=n1, normal2=n2, status=status,
                iters=it)


def shape_rays(ss, poses, idx, origins, dirs, max_toi, solid, nthreads=1, iters=False):
    """RayCast::toi_and_normal_with_ray of shape idx[k] against ray k; miss: toi -1, feature UINT32_MAX.
    iters: also return each ray's count of support evaluations (0 for the closed forms)."""
    L = lib()
    R = ss.R
    poses = np.ascontiguousarray(poses, dtype=R).reshape(-1, 7)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    o = np.ascontiguousarray(origins, dtype=R).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, dtype=R).reshape(-1, 3)
    n = len(idx)
    toi = np.zeros(n, R)
    nrm = np.zeros((n, 3), R)
    feat = np.zeros(n, np.uint32)
    it = np.zeros(n, np.int32) if iters else None
    getattr(L, f"orc_shape_rays_{ss.tag}")(*ss.args(), _p(poses), _p(idx), _p(o), _p(d), n, max_toi, int(solid),
                                            _p(toi), _p(nrm), _p(feat), _p(it) if iters else None, nthreads)
    return (toi, nrm, feat, it) if iters else (toi, nrm, feat)


def world_rays(ss, poses, boxes6, origins, dirs, max_toi, mode, nthreads=8):
    """Brute-force world ray queries over proxies with stored boxes (n, 6) (f32 only).
    mode 0: first_interference_with_ray -> (handle, toi, normal) per ray;
    mode 1: interferences_with_ray with solid shape casts, mode 2: stored boxes only ->
    (counts per ray, handles, toi, normal), handles ascending per ray."""
    L = lib()
    assert ss.R == np.float32
    poses = np.ascontiguousarray(poses, np.float32).reshape(-1, 7)
    boxes6 = np.ascontiguousarray(boxes6, np.float32).reshape(-1, 6)
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    nr = len(o)
    counts = np.zeros(nr, np.int64)
    if mode == 0:
        h = np.zeros(nr, np.uint32); t = np.zeros(nr, np.float32); n = np.zeros((nr, 3), np.float32)
        L.orc_world_rays_f32(*ss.args(), _p(poses), _p(boxes6), _p(o), _p(d), nr, max_toi, 0, _p(h), _p(t), _p(n),
                             _p(counts), nr, nthreads)
        return h, t, n
    total = L.orc_world_rays_f32(*ss.args(), _p(poses), _p(boxes6), _p(o), _p(d), nr, max_toi, mode, None, None,
                                 None, _p(counts), 0, nthreads)
    h = np.zeros(max(total, 1), np.uint32); t = np.zeros(max(total, 1), np.float32)
    n = np.zeros((max(total, 1), 3), np.float32)
    L.orc_world_rays_f32(*ss.args(), _p(poses), _p(boxes6), _p(o), _p(d), nr, max_toi, mode, _p(h), _p(t), _p(n),
                         _p(counts), total, nthreads)
    return counts, h[:total], t[:total], n[:total]


def polyhedra_gjk_step(ss, poses, ia, ib, prediction, last_dirs, nthreads=8):
    """ConvexPolyhedronConvexPolyhedronManifoldGenerator's GJK part per pair (f32): contact from
    last_gjk_dir (zero rows = None) and the new last_gjk_dir (dir_set False: unchanged)."""
    L = lib()
    assert ss.R == np.float32
    poses = np.ascontiguousarray(poses, np.float32).reshape(-1, 7)
    ia = np.ascontiguousarray(ia, np.uint32)
    ib = np.ascontiguousarray(ib, np.uint32)
    n = len(ia)
    ld = np.ascontiguousarray(last_dirs, np.float32).reshape(n, 3)
    has = np.zeros(n, np.uint8); dset = np.zeros(n, np.uint8)
    w1, w2, nr, nd = (np.zeros((n, 3), np.float32) for _ in range(4))
    depth = np.zeros(n, np.float32)
    L.orc_polyhedra_step_f32(*ss.args(), _p(poses), _p(ia), _p(ib), n, prediction, _p(ld), _p(has), _p(w1), _p(w2),
                             _p(nr), _p(depth), _p(nd), _p(dset), nthreads)
    return dict(has=has.astype(bool), world1=w1, world2=w2, normal=nr, depth=depth, dir=nd, dir_set=dset.astype(bool))


def hull_polyhedron(points, tris):
    """ConvexHull::try_new's faces, edges and adjacency (f32) from a hull's outward triangles;
    None where try_new returns None."""
    L = lib()
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    tr = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    n, nt = len(pts), len(tr)
    fn = np.zeros((nt, 3), np.float32)
    ff, fc = np.zeros(nt, np.uint32), np.zeros(nt, np.uint32)
    lv, le = np.zeros(3 * nt, np.uint32), np.zeros(3 * nt, np.uint32)
    ev, ed = np.zeros((3 * nt, 2), np.uint32), np.zeros((3 * nt, 3), np.float32)
    vf, vc = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    af, ae = np.zeros(3 * nt, np.uint32), np.zeros(3 * nt, np.uint32)
    nf, nl, ne = I64(0), I64(0), I64(0)
    rc = L.orc_hull_polyhedron_f32(_p(pts), n, _p(tr), nt, _p(fn), _p(ff), _p(fc), C.byref(nf), _p(lv), _p(le),
                                   C.byref(nl), _p(ev), _p(ed), C.byref(ne), _p(vf), _p(vc), _p(af), _p(ae))
    if rc != 0:
        return None
    nf, nl, ne = nf.value, nl.value, ne.value
    na = int(vc.sum())
    return dict(face_normals=fn[:nf], face_first=ff[:nf], face_count=fc[:nf], loop_vertex=lv[:nl], loop_edge=le[:nl],
                edge_vertices=ev[:ne], edge_dirs=ed[:ne], vertex_first=vf, vertex_count=vc, adj_face=af[:na],
                adj_edge=ae[:na])


class ManifoldStore:
    """Per-interaction state of the contact-manifold generators (last_gjk_dir and the
    ContactManifold's tracking cache), one slot per interaction, kept across steps."""

    def __init__(self, nslots):
        self._h = lib().orc_manifold_store_new(nslots)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().orc_manifold_store_free(self._h)
                self._h = None
        except Exception:  # interpreter shutdown
            pass

    def reset(self, slots):
        slots = np.ascontiguousarray(slots, np.uint32)
        assert lib().orc_manifold_reset(self._h, _p(slots), len(slots)) == 0

    def step(self, ss, hull_tris, tri_off, tri_cnt, poses, ia, ib, slots, linear, angular1, angular2, maxc=32,
             nthreads=8):
        """One NarrowPhase::update_contact per pair (f32): the manifold's contacts after the
        update, in cache order, at most maxc per pair."""
        assert ss.R == np.float32
        poses = np.ascontiguousarray(poses, np.float32).reshape(-1, 7)
        ia = np.ascontiguousarray(ia, np.uint32)
        ib = np.ascontiguousarray(ib, np.uint32)
        slots = np.ascontiguousarray(slots, np.uint32)
        tris = np.ascontiguousarray(hull_tris, np.uint32).reshape(-1)
        if len(tris) == 0:
            tris = np.zeros(3, np.uint32)
        toff = np.ascontiguousarray(tri_off, np.uint32)
        tcnt = np.ascontiguousarray(tri_cnt, np.uint32)
        n = len(ia)
        cnt = np.zeros(n, np.uint32)
        w1, w2, nr = (np.zeros((n, maxc, 3), np.float32) for _ in range(3))
        depth = np.zeros((n, maxc), np.float32)
        f1, f2 = np.zeros((n, maxc), np.uint32), np.zeros((n, maxc), np.uint32)
        dirs = np.zeros((n, 3), np.float32)
        rc = lib().orc_manifold_step_f32(self._h, *ss.args(), _p(tris), _p(toff), _p(tcnt), _p(poses), _p(ia), _p(ib),
                                         _p(slots), n, linear, angular1, angular2, maxc, _p(cnt), _p(w1), _p(w2),
                                         _p(nr), _p(depth), _p(f1), _p(f2), _p(dirs), nthreads)
        assert rc in (0, -2), f"manifold step failed ({rc})"
        return dict(count=cnt, world1=w1, world2=w2, normal=nr, depth=depth, feature1=f1, feature2=f2, dirs=dirs,
                    overflow=rc == -2)


def _mesh_args(ss, points, indices, mesh_pose7):
    R = ss.R
    pts = np.asarray(points, R).reshape(-1, 3)
    idx = np.asarray(indices, np.int64).reshape(-1, 3)
    tri = np.ascontiguousarray(pts[idx].reshape(-1, 9))
    mp = np.ascontiguousarray(mesh_pose7, R).reshape(1, 7)
    return tri, len(idx), mp


def mesh_contact(ss, points, indices, mesh_pose7, poses, shape_idx, prediction, nthreads=8):
    """contact_composite_shape_shape(mesh at mesh_pose7, shape shape_idx[k]) per query."""
    L = lib()
    R = ss.R
    tri, ntri, mp = _mesh_args(ss, points, indices, mesh_pose7)
    poses = np.ascontiguousarray(poses, R).reshape(-1, 7)
    qi = np.ascontiguousarray(shape_idx, np.uint32)
    n = len(qi)
    has = np.zeros(n, np.uint8); face = np.zeros(n, np.uint32)
    w1, w2, nr = (np.zeros((n, 3), R) for _ in range(3))
    depth = np.zeros(n, R)
    getattr(L, f"orc_mesh_queries_{ss.tag}")(_p(tri), ntri, _p(mp), *ss.args(), _p(poses), _p(qi), n, prediction, 0,
                                              _p(has), _p(w1), _p(w2), _p(nr), _p(depth), _p(face), None, nthreads)
    return dict(has=has.astype(bool), world1=w1, world2=w2, normal=nr, depth=depth, face=face)


def mesh_distance(ss, points, indices, mesh_pose7, poses, shape_idx, nthreads=8):
    """distance_composite_shape_shape(mesh at mesh_pose7, shape shape_idx[k]) per query."""
    L = lib()
    R = ss.R
    tri, ntri, mp = _mesh_args(ss, points, indices, mesh_pose7)
    poses = np.ascontiguousarray(poses, R).reshape(-1, 7)
    qi = np.ascontiguousarray(shape_idx, np.uint32)
    n = len(qi)
    dist = np.zeros(n, R)
    getattr(L, f"orc_mesh_queries_{ss.tag}")(_p(tri), ntri, _p(mp), *ss.args(), _p(poses), _p(qi), n, 0.0, 1, None,
                                              None, None, None, None, None, _p(dist), nthreads)
    return dist
