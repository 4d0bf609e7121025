"""GPU parity of the composite-shape queries (a TriMesh against convex shapes:
ncx_mesh_contact_batch, ncx_mesh_distance_batch through the C ABI) against the oracle's
contact_composite_shape_shape / distance_composite_shape_shape on the same seeded inputs, bit for
bit."""
import numpy as np
import pytest
import torch

import ncollide_amd as ncx
from ncollide_amd import workloads
from ncollide_amd.math import to_oracle_poses
from tests.test_gpu_parity import oracle_shapes

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _qrot(q, v):
    qv = q[:3]
    t = 2.0 * np.cross(qv, v)
    return v + q[3] * t + np.cross(qv, t)


def _scene(seed, n, cells=40, spacing=0.1):
    """Shapes of the scene distribution placed over a heightfield, near its surface (some under,
    some above), with the mesh and the shapes moved by one rigid pose."""
    pts, tri = workloads.heightfield_mesh(seed=seed, cells=cells, spacing=spacing)
    w = workloads.full_pipeline(seed=seed + 1, n=n, box=10.0, nclusters=1)
    rng = np.random.default_rng(seed + 2)
    ext = cells * spacing
    xy = rng.uniform(0.1 * ext, 0.9 * ext, (n, 2))
    # height of the nearest grid vertex, then an offset from 0.35 under to 0.6 over it
    ij = np.clip(np.rint(xy / spacing).astype(int), 0, cells)
    z = pts[ij[:, 1] * (cells + 1) + ij[:, 0], 2] + rng.uniform(-0.35, 0.6, n)
    local = np.concatenate([xy, z[:, None]], 1)
    ang = 0.3
    mq = np.array([0.0, np.sin(ang / 2), 0.0, np.cos(ang / 2)])
    mt = np.array([1.5, -0.7, 0.4])
    poses = w["poses"].copy()
    poses[:, :3] = np.array([_qrot(mq, p) for p in local]) + mt
    mpose8 = np.array([mt[0], mt[1], mt[2], 0.0, mq[0], mq[1], mq[2], mq[3]], np.float32)
    return pts, tri, w, poses.astype(np.float32), mpose8


@pytest.mark.parametrize("prediction", [0.0, 0.08])
def test_mesh_contacts_match_oracle(ctx, orc, prediction):
    pts, tri, w, poses, mpose = _scene(91, 3000)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    mesh = ncx.TriMesh(pts, tri, ctx=ctx)
    idx = np.arange(len(poses), dtype=np.uint32)
    g = mesh.contact_with_shapes(ss, poses, idx, prediction, pose=mpose)
    oss = oracle_shapes(orc, w["records"], w["pool"])
    mp7 = to_oracle_poses(mpose.reshape(1, 8))[0]
    r = orc.mesh_contact(oss, pts, tri, mp7, to_oracle_poses(poses), idx, np.float32(prediction))
    has = g["has_contact"].cpu().numpy()
    assert np.array_equal(has, r["has"]), f"has differs at {np.nonzero(has != r['has'])[0][:10]}"
    for key in ("depth", "normal", "world1", "world2"):
        assert np.array_equal(_bits(g[key].cpu().numpy()[has]), _bits(r[key][has])), f"{key} differs"
    assert np.array_equal(g["face"].cpu().numpy().view(np.uint32)[has], r["face"][has])
    types = w["records"]["type"]
    print(f"mesh contacts (prediction {prediction}): {has.sum()} of {len(has)}; by type "
          + ", ".join(f"{k}: {np.sum(has & (types == k))}" for k in range(4)))
    for k in range(4):
        assert np.sum(has & (types == k)) > 50
    # a second call with another set of queries (the combined shape set is refreshed per call)
    sel = idx[::3]
    g2 = mesh.contact_with_shapes(ss, poses, sel, prediction, pose=mpose)
    assert np.array_equal(_bits(g2["depth"].cpu().numpy()), _bits(g["depth"].cpu().numpy()[sel]))


def test_mesh_distances_match_oracle(ctx, orc):
    pts, tri, w, poses, mpose = _scene(92, 2500)
    poses[:, 2] += 0.3  # lift the shapes a little: most are apart
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    mesh = ncx.TriMesh(pts, tri, ctx=ctx)
    idx = np.arange(len(poses), dtype=np.uint32)
    g = mesh.distance_to_shapes(ss, poses, idx, pose=mpose).cpu().numpy()
    oss = oracle_shapes(orc, w["records"], w["pool"])
    mp7 = to_oracle_poses(mpose.reshape(1, 8))[0]
    r = orc.mesh_distance(oss, pts, tri, mp7, to_oracle_poses(poses), idx)
    assert np.array_equal(_bits(g), _bits(r)), f"distance differs at {np.nonzero(_bits(g) != _bits(r))[0][:10]}"
    print(f"mesh distances: {len(g)} queries, {np.sum(g > 0)} apart, median {np.median(g):.3f}")
    assert np.sum(g > 0) > 500 and np.sum(g == 0) > 50
