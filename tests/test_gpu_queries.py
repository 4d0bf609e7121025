"""GPU parity of the time-of-impact and shape ray-cast kernels (ncx_time_of_impact_batch,
ncx_shapes_cast_rays, called through the C ABI) against the f32 oracle on the same seeded inputs.
Both restate the reference operation by operation in f32, so every output is compared bit for
bit."""
import numpy as np
import pytest
import torch

import ncollide_amd as ncx
from ncollide_amd import query, workloads
from ncollide_amd.math import to_oracle_poses
from tests.test_gpu_parity import oracle_shapes

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _mixed_pairs(seed, n, npairs, box, nclusters):
    w = workloads.full_pipeline(seed=seed, n=n, box=box, nclusters=nclusters)
    rng = np.random.default_rng(seed + 100)
    ia = rng.integers(0, n, npairs).astype(np.uint32)
    ib = rng.integers(0, n, npairs).astype(np.uint32)
    keep = ia != ib
    return w, ia[keep], ib[keep], rng


@pytest.mark.parametrize("target", [0.0, 0.05])
def test_toi_mixed_shapes_match_oracle(ctx, orc, target):
    w, ia, ib, rng = _mixed_pairs(15, 3000, 40000, 12.0, 2)
    n = len(w["records"])
    # velocities: each shape heads toward a random point of the scene, speeds up to 4 per unit time
    t = w["poses"][:, :3]
    goal = t[rng.integers(0, n, n)]
    vel = ((goal - t) * rng.uniform(0.0, 0.5, (n, 1))).astype(np.float32)
    vel[::7] = 0.0  # some resting shapes
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    g = query.time_of_impact(ss, w["poses"], vel, ia, ib, 4.0, target)
    oss = oracle_shapes(orc, w["records"], w["pool"])
    r = orc.time_of_impact(oss, to_oracle_poses(w["poses"]), vel, ia, ib, np.float32(4.0), np.float32(target),
                           nthreads=8)
    has = g["has"].cpu().numpy()
    assert np.array_equal(has, r["has"]), f"has differs at {np.nonzero(has != r['has'])[0][:10]}"
    h = r["has"]
    assert np.array_equal(_bits(g["toi"].cpu().numpy()[h]), _bits(r["toi"][h])), "toi differs"
    for key in ("witness1", "witness2", "normal1", "normal2"):
        assert np.array_equal(_bits(g[key].cpu().numpy()[h]), _bits(r[key][h])), f"{key} differs"
    assert np.array_equal(g["status"].cpu().numpy()[h], r["status"][h])
    st = r["status"][h]
    types = w["records"]["type"]
    print(f"toi target {target}: {len(ia)} pairs, {h.sum()} hits, penetrating {np.sum(st == 3)}, "
          f"ball-ball {np.sum(h & (types[ia] == 0) & (types[ib] == 0))}, mean support evaluations "
          f"{r['iters'][h].mean():.2f}")
    assert h.sum() > 300


def test_toi_hull_pairs_match_oracle(ctx, orc):
    w = workloads.hull_pairs(seed=23, npairs=6000)
    n = len(w["records"])
    rng = np.random.default_rng(5)
    vel = np.zeros((n, 3), np.float32)
    # shape b of each pair moves back along -u (toward shape a) with some sideways drift
    u = w["dirs"]
    vel[1::2] = (-u * rng.uniform(0.2, 1.0, (len(u), 1)) + rng.normal(0, 0.1, (len(u), 3))).astype(np.float32)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    g = query.time_of_impact(ss, w["poses"], vel, w["ia"], w["ib"], 10.0, 0.0)
    oss = oracle_shapes(orc, w["records"], w["pool"])
    r = orc.time_of_impact(oss, to_oracle_poses(w["poses"]), vel, w["ia"], w["ib"], np.float32(10.0), np.float32(0.0),
                           nthreads=8)
    h = r["has"]
    assert np.array_equal(g["has"].cpu().numpy(), h)
    for key in ("toi", "witness1", "witness2", "normal1", "normal2"):
        assert np.array_equal(_bits(g[key].cpu().numpy()[h]), _bits(r[key][h])), f"{key} differs"
    assert np.array_equal(g["status"].cpu().numpy()[h], r["status"][h])
    print(f"hull toi: {h.sum()} of {len(h)} hit, penetrating {np.sum(r['status'][h] == 3)}")
    assert np.sum(r["status"][h] == 3) > 1000 and np.sum(r["status"][h] == 1) > 1000


@pytest.mark.parametrize("nr", [1, 63, 65, 130, 1000])
def test_shape_ray_casts_ragged_batches(ctx, orc, nr):
    """Batches that end inside a 64-ray window, over mixed shape kinds, equal the oracle."""
    w = workloads.full_pipeline(seed=19, n=500, box=10.0, nclusters=2)
    n = len(w["records"])
    rng = np.random.default_rng(nr)
    idx = rng.integers(0, n, nr).astype(np.uint32)
    c = w["poses"][idx, :3].astype(np.float64)
    o = c + rng.normal(0, 2.0, (nr, 3))
    d = c + rng.uniform(-0.3, 0.3, (nr, 3)) - o
    o, d = o.astype(np.float32), d.astype(np.float32)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    gt, gn, gf = query.cast_rays(ss, w["poses"], idx, o, d, 50.0, True)
    oss = oracle_shapes(orc, w["records"], w["pool"])
    rt, rn, rf = orc.shape_rays(oss, to_oracle_poses(w["poses"]), idx, o, d, np.float32(50.0), True, nthreads=8)
    assert np.array_equal(gf.cpu().numpy().view(np.uint32), rf)
    assert np.array_equal(_bits(gt.cpu().numpy()), _bits(rt))
    assert np.array_equal(_bits(gn.cpu().numpy()), _bits(rn))
    if nr >= 130:
        assert (rf != 0xFFFFFFFF).sum() > nr // 4


@pytest.mark.parametrize("solid", [True, False])
def test_shape_ray_casts_match_oracle(ctx, orc, solid):
    w = workloads.full_pipeline(seed=16, n=4000, box=20.0, nclusters=3)
    n = len(w["records"])
    rng = np.random.default_rng(17)
    nr = 60000
    idx = rng.integers(0, n, nr).astype(np.uint32)
    c = w["poses"][idx, :3].astype(np.float64)
    o = c + rng.normal(0, 2.0, (nr, 3))
    inside = rng.random(nr) < 0.2
    o[inside] = c[inside] + rng.uniform(-0.03, 0.03, (inside.sum(), 3))
    d = (c + rng.uniform(-0.3, 0.3, (nr, 3)) - o)
    d[inside] = rng.normal(size=(inside.sum(), 3))
    d *= rng.uniform(0.5, 2.0, (nr, 1))
    d[::97] = rng.normal(size=(len(d[::97]), 3))  # rays in random directions (mostly misses)
    o, d = o.astype(np.float32), d.astype(np.float32)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    gt, gn, gf = query.cast_rays(ss, w["poses"], idx, o, d, 50.0, solid)
    gt, gn, gf = gt.cpu().numpy(), gn.cpu().numpy(), gf.cpu().numpy().view(np.uint32)
    oss = oracle_shapes(orc, w["records"], w["pool"])
    rt, rn, rf = orc.shape_rays(oss, to_oracle_poses(w["poses"]), idx, o, d, np.float32(50.0), solid, nthreads=8)
    assert np.array_equal(gf, rf), f"features differ at {np.nonzero(gf != rf)[0][:10]}"
    assert np.array_equal(_bits(gt), _bits(rt)), f"toi differs at {np.nonzero(_bits(gt) != _bits(rt))[0][:10]}"
    assert np.array_equal(_bits(gn), _bits(rn)), "normals differ"
    hit = rf != 0xFFFFFFFF
    types = w["records"]["type"][idx]
    print(f"shape rays solid={solid}: {nr} rays, hits {hit.mean():.3f}; by type "
          + ", ".join(f"{k}: {np.mean(hit[types == k]):.2f}" for k in range(4)))
    for k in range(4):
        assert np.sum(hit & (types == k)) > 500


def test_shape_rays_capsule_iteration_cap_fixture(ctx):
    """Capsule rays whose GJK ray cast stops at its 10000-iteration cap (None) and converging
    controls (tests/golden/capsule_ray_fixed_points.npz): the GPU gives the fixture's outputs bit
    for bit."""
    import os
    from ncollide_amd.shape import Capsule
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "capsule_ray_fixed_points.npz"))
    ss = ncx.ShapeSet(shapes=[Capsule(h, r) for h, r in zip(z["half_height"], z["radius"])], ctx=ctx)
    gt, gn, gf = query.cast_rays(ss, z["poses"], z["shape"], z["origin"], z["dir"], 50.0, True)
    assert np.array_equal(gf.cpu().numpy().view(np.uint32), z["feature"])
    assert np.array_equal(_bits(gt.cpu().numpy()), _bits(z["toi"]))
    assert np.array_equal(_bits(gn.cpu().numpy()), _bits(z["normal"]))
    assert np.all(z["feature"][z["capped"]] == 0xFFFFFFFF)
