"""CPU tests that pin the oracle's composite-shape queries (a TriMesh against a convex shape:
contact_composite_shape_shape, distance_composite_shape_shape) to what mathematics fixes: shapes
resting on or hovering over a flat mesh (depth and distance are the shape's lowest point, the
normal the mesh's), and the same queries without the tree's pruning (every triangle tried), which
the restated box filter and best-first pruning must not change."""
import numpy as np
import pytest

from tests.test_oracle import pose7, shapeset
from tests.test_oracle_queries import qrot, rand_hull, rand_quat


def flat_mesh(cells=8, size=4.0):
    g = np.arange(cells + 1) * size - cells * size / 2
    x, y = np.meshgrid(g, g, indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), np.zeros(x.size)], 1)
    tri = []
    for i in range(cells):
        for j in range(cells):
            p00, p10, p01, p11 = i * (cells + 1) + j, (i + 1) * (cells + 1) + j, i * (cells + 1) + j + 1, \
                (i + 1) * (cells + 1) + j + 1
            tri += [(p00, p10, p01), (p10, p11, p01)]
    return pts, np.array(tri, np.uint32)


def wavy_mesh(rng, cells=10, size=1.0):
    g = np.arange(cells + 1) * size - cells * size / 2
    x, y = np.meshgrid(g, g, indexing="ij")
    z = 0.3 * np.sin(x * 1.3) * np.cos(y * 0.9) + rng.uniform(-0.05, 0.05, x.shape)
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    tri = []
    for i in range(cells):
        for j in range(cells):
            p00, p10, p01, p11 = i * (cells + 1) + j, (i + 1) * (cells + 1) + j, i * (cells + 1) + j + 1, \
                (i + 1) * (cells + 1) + j + 1
            tri += [(p00, p10, p01), (p10, p11, p01)]
    return pts, np.array(tri, np.uint32)


def _shapes(rng, n):
    out = []
    for k in range(n):
        kind = k % 4
        if kind == 0:
            out.append(("ball", rng.uniform(0.2, 0.6)))
        elif kind == 1:
            out.append(("cuboid", rng.uniform(0.2, 0.6, 3)))
        elif kind == 2:
            out.append(("capsule", rng.uniform(0.1, 0.4), rng.uniform(0.1, 0.3)))
        else:
            out.append(("hull", rand_hull(rng, n=20, scale=(0.2, 0.6))))
    return out


def _lowest_z(shape, pose):
    """The least z over the posed shape (its support point along -z)."""
    t, q = pose[:3], pose[3:]
    if shape[0] == "ball":
        return t[2] - shape[1]
    if shape[0] == "cuboid":
        h = np.asarray(shape[1])
        c = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * h
        return min(qrot(q, v)[2] for v in c) + t[2]
    if shape[0] == "capsule":
        hh, r = shape[1], shape[2]
        return min(qrot(q, np.array([0, s * hh, 0]))[2] for s in (-1, 1)) + t[2] - r
    return min(qrot(q, v)[2] for v in shape[1]) + t[2]


def test_composite_on_flat_mesh_closed_forms(orc):
    rng = np.random.default_rng(81)
    pts, tri = flat_mesh()
    n = 80
    shapes = _shapes(rng, n)
    ss = shapeset(orc, shapes)
    poses = []
    for s in shapes:
        q = rand_quat(rng)
        p = pose7((rng.uniform(-10, 10), rng.uniform(-10, 10), 0.0), q)
        p[2] = -_lowest_z(s, p) + rng.uniform(-0.03, 0.3)  # lowest point between 0.03 under and 0.3 over
        poses.append(p)
    poses = np.array(poses)
    idx = np.arange(n)
    mp = pose7((0, 0, 0))
    d = orc.mesh_distance(ss, pts, tri, mp, poses, idx)
    c = orc.mesh_contact(ss, pts, tri, mp, poses, idx, 0.05)
    for k in range(n):
        low = _lowest_z(shapes[k], poses[k])
        kind = shapes[k][0]
        tol = 1e-6 if kind in ("ball", "capsule") else 1e-9
        if low > 0:
            assert abs(d[k] - low) <= tol + 1e-7 * low, (k, kind, d[k], low)
        else:
            assert d[k] == 0.0
        assert c["has"][k] == (low <= 0.05), (k, kind, low)
        if c["has"][k]:
            assert abs(c["depth"][k] + low) <= 1e-6, (k, kind, c["depth"][k], -low)
            assert np.allclose(c["normal"][k], (0, 0, 1), atol=1e-5), (k, kind, c["normal"][k])
            assert abs(c["world1"][k][2]) <= 1e-6  # the mesh's witness lies on the mesh


def test_composite_pruning_changes_nothing(orc):
    """Box filter (contact) and best-first pruning (distance) against every triangle tried."""
    rng = np.random.default_rng(82)
    pts, tri = wavy_mesh(rng)
    n = 60
    shapes = _shapes(rng, n)
    ss = shapeset(orc, shapes)
    mq = rand_quat(rng)
    mp = pose7((0.3, -0.2, 0.1), mq)
    poses = []
    for s in shapes:
        local = np.array([rng.uniform(-4, 4), rng.uniform(-4, 4), rng.uniform(-0.3, 0.9)])
        poses.append(pose7(qrot(mq, local) + mp[:3], rand_quat(rng)))
    poses = np.array(poses)
    idx = np.arange(n)
    d = orc.mesh_distance(ss, pts, tri, mp, poses, idx)
    c = orc.mesh_contact(ss, pts, tri, mp, poses, idx, 0.1)
    # every triangle as its own one-triangle mesh
    best_d = np.full(n, np.inf)
    best_c = [None] * n
    for f in range(len(tri)):
        t1 = tri[f:f + 1]
        df = orc.mesh_distance(ss, pts, t1, mp, poses, idx)
        best_d = np.minimum(best_d, df)
        cf = orc.mesh_contact(ss, pts, t1, mp, poses, idx, 0.1)
        for k in np.nonzero(cf["has"])[0]:
            if best_c[k] is None or cf["depth"][k] > best_c[k][0]:
                best_c[k] = (cf["depth"][k], f, cf["normal"][k])
    assert np.array_equal(d, best_d)
    for k in range(n):
        assert c["has"][k] == (best_c[k] is not None), k
        if c["has"][k]:
            assert c["depth"][k] == best_c[k][0] and c["face"][k] == best_c[k][1], k
    print(f"composite: {np.sum(c['has'])} contacts, {np.sum(d == 0)} touching of {n}")
    assert np.sum(c["has"]) > 10 and np.sum(d > 0) > 10
