"""GPU parity: the HIP kernels of libncx.so (called through the C ABI) against the
oracle on the same seeded inputs.

Bars: pair sets, ray-hit features and AABBs bit-exact; contact/distance within
1e-5 relative + 1e-6 absolute, normals within 1e-4 rad (the tolerance of the
north-star contract), and the fraction of bit-identical results is reported.
"""
import numpy as np
import pytest
import torch

import ncollide_amd as ncx
from ncollide_amd import workloads
from ncollide_amd.math import to_oracle_poses

pytestmark = pytest.mark.gpu

REL, ABS, ANG = 1e-5, 1e-6, 1e-4


def oracle_shapes(orc, records, pool, dtype=np.float32):
    params = np.stack([records["p0"], records["p1"], records["p2"]], 1)
    return orc.ShapeSet(records["type"], params, records["hull_offset"], records["hull_count"], pool, dtype=dtype)


def assert_contacts_match(gpu, ref, label):
    has_g = gpu["has_contact"].cpu().numpy()
    has_r = ref["has"]
    mism = np.nonzero(has_g != has_r)[0]
    assert len(mism) == 0, f"{label}: has_contact differs at {mism[:10]}"
    idx = np.nonzero(has_r)[0]
    dg = gpu["depth"].cpu().numpy()[idx].astype(np.float64)
    dr = ref["depth"][idx].astype(np.float64)
    err = np.abs(dg - dr)
    ok = err <= REL * np.abs(dr) + ABS
    assert ok.all(), f"{label}: depth off at {idx[~ok][:10]}: gpu {dg[~ok][:5]} ref {dr[~ok][:5]}"
    ng = gpu["normal"].cpu().numpy()[idx].astype(np.float64)
    nr = ref["normal"][idx].astype(np.float64)
    cosang = np.clip(np.sum(ng * nr, 1) / (np.linalg.norm(ng, axis=1) * np.linalg.norm(nr, axis=1) + 1e-300), -1, 1)
    ang = np.arccos(cosang)
    assert np.all(ang <= ANG), f"{label}: normal off by {ang.max()} rad"
    exact = np.mean(dg == dr) if len(idx) else 1.0
    print(f"{label}: {len(idx)} contacts, bit-exact depth fraction {exact:.4f}, max |d depth| {err.max() if len(err) else 0:.3g}")
    return exact


# ------------------------------------------------------------------ ray casting
def test_raycast_heightfield_matches_oracle(ctx, orc):
    pts, tri = workloads.heightfield_mesh(seed=2)
    o, d = workloads.heightfield_rays(seed=2, n=20000)
    mesh = ncx.TriMesh(pts, tri, ctx=ctx)
    toi, nrm, feat = mesh.toi_and_normal_with_ray(o, d, 100.0)
    toi, nrm, feat = toi.cpu().numpy(), nrm.cpu().numpy(), feat.cpu().numpy().view(np.uint32)
    om = orc.TriMesh(pts, tri)
    sel = np.arange(0, 20000, 4)
    bt, bn, bf, nh = om.brute(o[sel], d[sel], 100.0, nthreads=8)
    assert np.array_equal(feat[sel], bf), "feature ids differ from the brute-force oracle"
    assert np.array_equal(toi[sel], bt)
    assert np.array_equal(nrm[sel], bn)
    rt, rn, rf, _ = om.cast(o, d, 100.0, nthreads=8)  # the reference's best-first search over its BVT
    diff = np.nonzero(feat != rf)[0]
    for k in diff:  # only exact toi ties may resolve to another triangle
        assert toi[k] == rt[k]
    assert np.array_equal(toi, rt)
    print(f"rays: {len(o)}, hits {np.mean(feat != 0xFFFFFFFF):.3f}, tie-resolved differences {len(diff)}")


def test_raycast_axis_aligned_rays_on_grid_lines(ctx, orc):
    # Rays with zero direction components through a many-triangle mesh (the traversal's boxes
    # take the containment branch of the slab test), origins exactly on vertex coordinates and
    # cell edges, where box faces lie on the ray.
    pts, tri = workloads.heightfield_mesh(seed=4, cells=32)
    rng = np.random.default_rng(7)
    xs = np.unique(pts[:, 0])
    n = 3000
    o = np.empty((n, 3), np.float32)
    d = np.zeros((n, 3), np.float32)
    o[:, 0] = rng.choice(xs, n)
    o[:, 1] = rng.choice(xs, n)
    o[:, 2] = 3.0
    half = n // 3
    o[half:2 * half, 0] += rng.uniform(0, 0.1, half).astype(np.float32)  # on a grid line in y only
    d[:2 * half, 2] = -1.0                                              # straight down
    k = n - 2 * half
    axis = rng.integers(0, 2, k)                                        # along x or y, from outside
    sign = rng.choice([-1.0, 1.0], k)
    o[2 * half:, 2] = rng.uniform(-0.6, 0.6, k)
    o[2 * half:, 0] = np.where(axis == 0, np.where(sign > 0, -1.0, 4.2), o[2 * half:, 0])
    o[2 * half:, 1] = np.where(axis == 1, np.where(sign > 0, -1.0, 4.2), o[2 * half:, 1])
    d[2 * half:, 0] = np.where(axis == 0, sign, 0.0)
    d[2 * half:, 1] = np.where(axis == 1, sign, 0.0)
    mesh = ncx.TriMesh(pts, tri, ctx=ctx)
    toi, nrm, feat = mesh.toi_and_normal_with_ray(o, d, 100.0)
    toi, nrm, feat = toi.cpu().numpy(), nrm.cpu().numpy(), feat.cpu().numpy().view(np.uint32)
    om = orc.TriMesh(pts, tri)
    bt, bn, bf, _ = om.brute(o, d, 100.0, nthreads=8)
    assert np.array_equal(toi, bt), "toi differs from the least hit over all triangles"
    assert np.array_equal(feat, bf)
    assert np.array_equal(nrm, bn)
    # The reference's best-first search (bvh.rs:101-160) keeps the first hit found and stops once
    # no box can hold a smaller cost.  A ray through a mesh vertex hits the triangles around it
    # at parameters a rounding apart, and the search may stop on one 1-2 ulps past the least
    # (its box's entry parameter already exceeds the better hit): only there may it differ.
    rt, rn, rf, _ = om.cast(o, d, 100.0, nthreads=8)
    diff = np.nonzero(toi != rt)[0]
    ulps = rt[diff].view(np.int32).astype(np.int64) - toi[diff].view(np.int32).astype(np.int64)
    assert np.all((ulps >= 1) & (ulps <= 2)), ulps
    assert np.all(diff < half), "a reference difference away from a mesh vertex"
    print(f"rays {n}, hits {np.mean(feat != 0xFFFFFFFF):.3f}, reference search 1-2 ulps farther on {len(diff)}")


def test_raycast_edge_cases(ctx, orc):
    # single triangle, rays along axes (zero direction components), back faces, max_toi cut
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    tri = np.array([[0, 1, 2]], np.uint32)
    o = np.array([[0.25, 0.25, 1], [0.25, 0.25, -1], [0.25, 0.25, 1], [2, 2, 1], [0.0, 0.0, 1], [0.25, 0.25, 5]],
                 np.float32)
    d = np.array([[0, 0, -1], [0, 0, 1], [0, 0, 1], [0, 0, -1], [0, 0, -1], [0, 0, -1]], np.float32)
    mesh = ncx.TriMesh(pts, tri, ctx=ctx)
    toi, nrm, feat = mesh.toi_and_normal_with_ray(o, d, 2.0)
    om = orc.TriMesh(pts, tri)
    rt, rn, rf, _ = om.cast(o, d, 2.0)
    assert np.array_equal(feat.cpu().numpy().view(np.uint32), rf)
    assert np.array_equal(toi.cpu().numpy(), rt)
    f = feat.cpu().numpy().view(np.uint32)
    assert f[0] == 0 and f[1] == 1 and f[2] == 0xFFFFFFFF and f[3] == 0xFFFFFFFF and f[5] == 0xFFFFFFFF


@pytest.mark.parametrize("layout", ["geometric", "coincident"])
def test_raycast_unbalanced_meshes(ctx, orc, layout):
    """Meshes that a surface-area split would cut one triangle at a time (3,000 small triangles
    at geometrically growing distances along x), or cannot split at all (2,000 copies of one
    triangle): the tree stays within the traversal stacks and every hit equals brute force."""
    rng = np.random.default_rng(11)
    if layout == "geometric":
        n = 3000
        x = (1.02 ** np.arange(n)).astype(np.float32)
        base = np.stack([x, np.zeros(n), np.zeros(n)], 1)
        pts = np.concatenate([base, base + [0, 1, 0], base + [0, 0, 1]]).astype(np.float32)
        tri = np.stack([np.arange(n), np.arange(n) + n, np.arange(n) + 2 * n], 1).astype(np.uint32)
        # rays along +x from between the triangles, through their (y, z) extent
        o = np.stack([x[rng.integers(0, n, 2000)] * rng.uniform(0.5, 1.0, 2000), rng.uniform(0.05, 0.45, 2000),
                      rng.uniform(0.05, 0.45, 2000)], 1).astype(np.float32)
        d = np.tile(np.array([[1, 0, 0]], np.float32), (2000, 1))
    else:
        n = 2000
        pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
        tri = np.tile(np.array([[0, 1, 2]], np.uint32), (n, 1))
        o = np.concatenate([rng.uniform(-0.2, 1.2, (1000, 2)), np.ones((1000, 1))], 1).astype(np.float32)
        d = np.tile(np.array([[0, 0, -1]], np.float32), (1000, 1))
    mesh = ncx.TriMesh(pts, tri, ctx=ctx)
    toi, nrm, feat = mesh.toi_and_normal_with_ray(o, d, 1e30)
    om = orc.TriMesh(pts, tri)
    bt, bn, bf, _ = om.brute(o, d, 1e30, nthreads=8)
    assert np.array_equal(feat.cpu().numpy().view(np.uint32), bf)
    assert np.array_equal(toi.cpu().numpy(), bt)
    assert np.array_equal(nrm.cpu().numpy(), bn)
    assert (bf != 0xFFFFFFFF).mean() > 0.2


@pytest.mark.parametrize("ntri", [2, 3, 5, 9])
def test_raycast_small_meshes(ctx, orc, ntri):
    """Meshes of a few triangles, where the cast tree's leaves of up to two triangles meet its
    root: every hit equals brute force."""
    rng = np.random.default_rng(ntri)
    pts = rng.uniform(-1, 1, (3 * ntri, 3)).astype(np.float32)
    tri = np.arange(3 * ntri, dtype=np.uint32).reshape(ntri, 3)
    o = rng.uniform(-2, 2, (3000, 3)).astype(np.float32)
    d = workloads.unit_vectors(rng, 3000).astype(np.float32)
    d = np.where((o @ np.ones(3))[:, None] > 0, -np.abs(d), np.abs(d)).astype(np.float32)  # mostly inward
    mesh = ncx.TriMesh(pts, tri, ctx=ctx)
    toi, nrm, feat = mesh.toi_and_normal_with_ray(o, d, 10.0)
    om = orc.TriMesh(pts, tri)
    bt, bn, bf, _ = om.brute(o, d, 10.0, nthreads=8)
    assert np.array_equal(feat.cpu().numpy().view(np.uint32), bf)
    assert np.array_equal(toi.cpu().numpy(), bt)
    assert np.array_equal(nrm.cpu().numpy(), bn)
    assert (bf != 0xFFFFFFFF).sum() > 50


def test_raycast_posed_mesh(ctx, orc):
    pts, tri = workloads.heightfield_mesh(seed=7, cells=64)
    rng = np.random.default_rng(5)
    q = workloads.uniform_quats(rng, 1)[0]
    t = np.array([1.0, -2.0, 0.5], np.float32)
    pose8 = np.zeros(8, np.float32)
    pose8[:3] = t
    pose8[4:] = q
    o = rng.uniform(-5, 5, (4000, 3)).astype(np.float32)
    d = workloads.unit_vectors(rng, 4000).astype(np.float32)
    mesh = ncx.TriMesh(pts, tri, ctx=ctx)
    toi, nrm, feat = mesh.toi_and_normal_with_ray(o, d, 50.0, pose=pose8)
    om = orc.TriMesh(pts, tri)
    bt, bn, bf, _ = om.brute(o, d, 50.0, pose=np.concatenate([t, q]), nthreads=8)
    assert np.array_equal(feat.cpu().numpy().view(np.uint32), bf)
    assert np.array_equal(toi.cpu().numpy(), bt)
    assert np.array_equal(nrm.cpu().numpy(), bn)


# ------------------------------------------------------------------ shape AABBs and broad phase
def test_shape_aabbs_bit_exact(ctx, orc):
    w = workloads.full_pipeline(seed=5, n=4096)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    bb = ss.aabbs(poses, margin=0.0).cpu().numpy().T
    ref = orc.shape_aabbs(oracle_shapes(orc, w["records"], w["pool"]), to_oracle_poses(w["poses"]))
    assert np.array_equal(bb, ref)


def test_shape_aabbs_of_large_hulls_bit_exact(ctx, orc):
    # hulls of up to 256 points: a wavefront's 64 lanes take several passes over each
    w = workloads.hull_pairs(seed=9, npairs=1024)
    assert w["records"]["hull_count"].max() > 128
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    bb = ss.aabbs(poses, margin=0.01).cpu().numpy().T
    ref = orc.shape_aabbs(oracle_shapes(orc, w["records"], w["pool"]), to_oracle_poses(w["poses"]))
    ref = ref + np.array([-0.01, -0.01, -0.01, 0.01, 0.01, 0.01], np.float32)
    assert np.array_equal(bb, ref)


def test_broad_phase_plain_pipeline(ctx, orc):
    w = workloads.plain_pipeline(seed=1)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    bb = ss.aabbs(poses, margin=w["margin"]).cpu().numpy().T
    bp = ncx.DBVTBroadPhase(w["margin"], ctx=ctx)
    hs = bp.create_proxies(bb)
    assert np.array_equal(hs, np.arange(1024))
    bp.update()
    ref = orc.BroadPhase(w["margin"])
    for b in bb:
        ref.create_proxy(b)
    ref.update()
    assert np.array_equal(bp.pairs(), ref.pairs())
    assert np.array_equal(bp.pairs(), orc.pairs_brute(bb))
    s, t = bp.events()
    rs, rt = ref.events()
    assert np.array_equal(s, rs) and len(t) == 0


def test_broad_phase_dbvt_example(ctx):
    """examples/dbvt_broad_phase3d.rs:21-61."""
    bp = ncx.DBVTBroadPhase(0.2, ctx=ctx)
    hs = bp.create_proxies(np.array([[x - .5, y - .5, -.5, x + .5, y + .5, .5]
                                     for x, y in [(0, 0), (0, .5), (.5, 0), (.5, .5)]], np.float32))
    bp.update()
    assert bp.num_interferences() == 6
    bp.remove(hs[:2])
    bp.update()
    assert bp.num_interferences() == 1


def test_broad_phase_moving_matches_reference(ctx, orc):
    """Several steps of deferred_set_bounding_volume + update, with removals and re-created proxies,
    against the DBVTBroadPhase restatement (pair sets and started/stopped events)."""
    rng = np.random.default_rng(12)
    n = 3000
    c = rng.uniform(0, 30, (n, 3)).astype(np.float32)
    r = rng.uniform(0.2, 0.8, (n, 1)).astype(np.float32)
    bb = np.concatenate([c - r, c + r], 1)
    bp = ncx.DBVTBroadPhase(0.05, ctx=ctx)
    ref = orc.BroadPhase(0.05)
    hs = bp.create_proxies(bb)
    rh = [ref.create_proxy(b) for b in bb]
    assert np.array_equal(hs, rh)
    for step in range(6):
        bp.update()
        ref.update()
        assert np.array_equal(bp.pairs(), ref.pairs()), f"step {step}"
        s, t = bp.events()
        rs, rt = ref.events()
        assert np.array_equal(s, rs) and np.array_equal(t, rt), f"events step {step}"
        c = c + rng.normal(0, 0.1, c.shape).astype(np.float32)
        bb = np.concatenate([c - r, c + r], 1)
        alive = np.array(sorted(set(range(n))), np.uint32)
        bp.deferred_set_bounding_volumes(alive, bb)
        for h in alive:
            ref.deferred_set_bounding_volume(int(h), bb[h])


def test_broad_phase_clustered(ctx, orc):
    w = workloads.full_pipeline(seed=5, n=20000)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    bb = ss.aabbs(torch.as_tensor(w["poses"], device=ctx.device), margin=0.02).cpu().numpy().T
    bp = ncx.DBVTBroadPhase(0.02, ctx=ctx)
    bp.create_proxies(bb)
    bp.update()
    assert np.array_equal(bp.pairs(), orc.pairs_brute(bb))


# ------------------------------------------------------------------ contact / distance
def test_contact_plain_pipeline(ctx, orc):
    w = workloads.plain_pipeline(seed=1)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    bb = ss.aabbs(poses, margin=w["margin"]).cpu().numpy().T
    bp = ncx.DBVTBroadPhase(w["margin"], ctx=ctx)
    bp.create_proxies(bb)
    bp.update()
    keys = bp.pairs()
    ia = (keys >> np.uint64(32)).astype(np.uint32)
    ib = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    g = ncx.query.contact(ss, poses, ia, ib, w["prediction"])
    r = orc.contact(oracle_shapes(orc, w["records"], w["pool"]), to_oracle_poses(w["poses"]), ia, ib,
                    w["prediction"])
    assert_contacts_match(g, r, "plain pipeline")


def test_contact_hull_pairs(ctx, orc):
    w = workloads.hull_pairs(seed=3, npairs=2048)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    g = ncx.query.contact(ss, poses, w["ia"], w["ib"], 0.0)
    r = orc.contact(oracle_shapes(orc, w["records"], w["pool"]), to_oracle_poses(w["poses"]), w["ia"], w["ib"], 0.0,
                    nthreads=8)
    assert_contacts_match(g, r, "hull pairs")
    assert np.array_equal(g["gjk_iters"].cpu().numpy(), r["gjk_iters"])


def test_contact_mixed_shapes(ctx, orc):
    w = workloads.full_pipeline(seed=5, n=3000, box=30.0, nclusters=4)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    bb = ss.aabbs(poses, margin=0.02).cpu().numpy().T
    keys = orc.pairs_brute(bb)
    ia = (keys >> np.uint64(32)).astype(np.uint32)
    ib = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    assert len(ia) > 500
    for pred in (0.0, 0.3):
        g = ncx.query.contact(ss, poses, ia, ib, pred)
        r = orc.contact(oracle_shapes(orc, w["records"], w["pool"]), to_oracle_poses(w["poses"]), ia, ib, pred,
                        nthreads=8)
        assert_contacts_match(g, r, f"mixed shapes prediction {pred}")


@pytest.mark.parametrize("variant", range(5))
def test_contact_every_first_tier_variant_matches_oracle(ctx, orc, variant):
    """Each first-tier kernel variant (lanes per pair, EPA arena size), pinned on every bucket:
    the same bit-exact contacts.  The 24-vertex arena sends the longer EPA runs to the second
    tier, so this also runs the T1 -> T2 hand-over on many pairs."""
    w = workloads.full_pipeline(seed=6, n=6000, box=30.0, nclusters=3)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    ss.set_tuning([variant] * 4)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    bb = ss.aabbs(poses, margin=0.02).cpu().numpy().T
    keys = orc.pairs_brute(bb)
    ia = (keys >> np.uint64(32)).astype(np.uint32)
    ib = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    g = ncx.query.contact(ss, poses, ia, ib, 0.0)
    r = orc.contact(oracle_shapes(orc, w["records"], w["pool"]), to_oracle_poses(w["poses"]), ia, ib, 0.0,
                    nthreads=8)
    assert_contacts_match(g, r, f"variant {variant}")
    assert ss.tuning("contact")[0] == [ss.VARIANTS[variant]] * 4
    print(f"variant {ss.VARIANTS[variant]}: {len(ia)} pairs, {int(r['has'].sum())} contacts, "
          f"{int(np.sum(r['epa_iters'] > 20))} EPA runs past 20 iterations")


@pytest.mark.parametrize("variant", range(5))
def test_distance_every_first_tier_variant_matches_oracle(ctx, orc, variant):
    """Each distance kernel variant (8, 16, 32, 4, 2 lanes per pair), pinned on every bucket:
    the same bit-exact distances and GJK iteration counts."""
    w = workloads.full_pipeline(seed=6, n=6000, box=30.0, nclusters=3)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    ss.set_tuning([variant] * 4, query="distance")
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    bb = ss.aabbs(poses, margin=0.3).cpu().numpy().T
    keys = orc.pairs_brute(bb)
    ia = (keys >> np.uint64(32)).astype(np.uint32)
    ib = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    d, it = ncx.query.distance(ss, poses, ia, ib)
    rd, rit = orc.distance(oracle_shapes(orc, w["records"], w["pool"]), to_oracle_poses(w["poses"]), ia, ib,
                           nthreads=8)
    assert np.array_equal(d.cpu().numpy().view(np.uint32), np.asarray(rd, np.float32).view(np.uint32))
    assert np.array_equal(it.cpu().numpy(), np.asarray(rit))
    assert ss.tuning("distance")[0] == [ss.DISTANCE_VARIANTS[variant]] * 4
    assert (np.asarray(rd) > 0).mean() > 0.3


def _curved_pairs(seed, n):
    """n isolated Ball-Capsule / Capsule-Capsule pairs at random overlaps (shapes 2i, 2i + 1)."""
    rng = np.random.default_rng(seed)
    rec = np.zeros(2 * n, workloads.SHAPE_DTYPE)
    kind = rng.integers(0, 2, n)
    rec["type"][0::2] = np.where(kind == 0, workloads.BALL, workloads.CAPSULE)
    rec["type"][1::2] = workloads.CAPSULE
    rec["p0"][0::2] = rng.uniform(0.1, 0.5, n)
    rec["p0"][1::2] = rng.uniform(0.1, 0.5, n)
    rec["p1"][1::2] = rng.uniform(0.05, 0.3, n)
    rec["p1"][0::2] = np.where(kind == 1, rng.uniform(0.05, 0.3, n), 0.0)
    t = np.zeros((2 * n, 3), np.float32)
    t[1::2] = (workloads.unit_vectors(rng, n) * rng.uniform(0.0, 0.6, (n, 1))).astype(np.float32)
    poses8 = workloads.poses8(t, workloads.uniform_quats(rng, 2 * n))
    return rec, poses8, rng


@pytest.mark.parametrize("seed", [11, 12])
def test_long_epa_runs_through_every_tier(ctx, orc, seed):
    # Curved surfaces make EPA run long (tens to ~2,000 iterations), so pairs finish in every
    # arena tier (44, 128, 512 and 1,536 vertices in LDS, then device memory), and runs that
    # outgrow a tier are saved and continued by the next.  Every output must be bit-identical to
    # the f32 oracle.  Seed 12 holds two runs of 1,600 and 2,142 iterations (device-memory tier).
    n = 400000
    rec, poses8, rng = _curved_pairs(seed, n)
    pool = np.zeros((1, 3), np.float32)
    ia = np.arange(0, 2 * n, 2, dtype=np.uint32)
    ib = ia + 1
    r = orc.contact(oracle_shapes(orc, rec, pool), to_oracle_poses(poses8), ia, ib, 0.0, nthreads=8)
    e = r["epa_iters"]
    tiers = [int(np.sum((e > a) & (e <= b))) for a, b in ((0, 40), (40, 124), (124, 508), (508, 1532), (1532, 1 << 30))]
    assert tiers[3] > 0 and tiers[2] > 0, tiers
    if seed == 12:
        assert tiers[4] == 2, tiers
    sel = np.nonzero(e > (40 if seed == 11 else 124))[0]
    sel = np.concatenate([sel, rng.choice(np.nonzero(e <= 40)[0], 20000, replace=False)])
    ss = ncx.ShapeSet(records=rec, pool=pool, ctx=ctx)
    g = ncx.query.contact(ss, torch.as_tensor(poses8, device=ctx.device), ia[sel], ib[sel], 0.0)
    assert np.array_equal(g["has_contact"].cpu().numpy(), r["has"][sel])
    for k in ("depth", "normal", "world1", "world2"):
        gv = g[k].cpu().numpy()
        rv = r[k][sel]
        bad = np.nonzero(np.any((gv != rv).reshape(len(sel), -1), 1))[0]
        assert len(bad) == 0, f"{k} differs on {len(bad)} pairs, EPA iterations {e[sel][bad][:10]}"
    print(f"seed {seed}: pairs by EPA iterations (<=40, <=124, <=508, <=1532, more): {tiers}, checked {len(sel)}, "
          f"max {e.max()}")


def test_distance_mixed_shapes(ctx, orc):
    w = workloads.full_pipeline(seed=6, n=2000, box=20.0, nclusters=2)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    rng = np.random.default_rng(1)
    ia = rng.integers(0, 2000, 4000).astype(np.uint32)
    ib = rng.integers(0, 2000, 4000).astype(np.uint32)
    keep = ia != ib
    ia, ib = ia[keep], ib[keep]
    d, it = ncx.query.distance(ss, poses, ia, ib)
    rd, rit = orc.distance(oracle_shapes(orc, w["records"], w["pool"]), to_oracle_poses(w["poses"]), ia, ib,
                           nthreads=8)
    d = d.cpu().numpy().astype(np.float64)
    err = np.abs(d - rd)
    assert np.all(err <= REL * np.abs(rd) + ABS), f"max err {err.max()}"
    print(f"distance: bit-exact fraction {np.mean(d == rd):.4f}")


def test_contact_known_answers_gpu(ctx, orc):
    """epa3.rs:7-22 and contact_query3d.rs in f32 on the GPU.  The reference test runs in f64
    (the oracle reproduces 0.5 / 1.8 exactly there, see test_oracle.py); in f32 the (0, 0.2, 0)
    case takes EPA's numerical-issue return (epa3.rs:397-402), so the GPU is held to the f32 oracle."""
    shapes = [ncx.Cuboid([2, 1, 1]), ncx.Cuboid([2, 1, 1]), ncx.Ball(1.0), ncx.Cuboid([1, 1, 1])]
    ss = ncx.ShapeSet(shapes, ctx=ctx)
    oss = oracle_shapes(orc, ss.records, ss.pool)
    poses = np.zeros((4, 8), np.float32)
    poses[:, 7] = 1
    poses[0, :3] = [3.5, 0, 0]
    g = ncx.query.contact(ss, poses, [0], [1], 10.0)
    assert g["depth"].item() == 0.5 and np.array_equal(g["normal"].cpu().numpy()[0], [-1, 0, 0])
    poses[0, :3] = [0, 0.2, 0]
    g = ncx.query.contact(ss, poses, [0], [1], 10.0)
    r = orc.contact(oss, to_oracle_poses(poses), [0], [1], 10.0)
    assert g["depth"].item() == r["depth"][0]
    assert np.array_equal(g["normal"].cpu().numpy()[0], r["normal"][0])
    for k in (1, 2, 3):
        poses[2, :3] = [k, k, k]
        g = ncx.query.contact(ss, poses, [2], [3], 1.0)
        if k == 3:
            assert not g["has_contact"].item()
        else:
            assert abs(g["depth"].item() - (1 - np.sqrt(3) * (k - 1))) < 1e-6


def test_contact_bad_index_raises(ctx):
    ss = ncx.ShapeSet([ncx.Ball(1.0), ncx.Ball(1.0)], ctx=ctx)
    poses = np.zeros((2, 8), np.float32)
    poses[:, 7] = 1
    with pytest.raises(ncx.NcxError):
        ncx.query.contact(ss, poses, [0], [5], 0.0)


def test_warm_started_contacts_match_oracle(ctx, orc):
    """Warm start with given init directions and a 32-iteration cap (plain GJK recomputes capped
    pairs) against the oracle run with the same directions and the same fallback rule."""
    w = workloads.hull_pairs(seed=13, npairs=1500)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    rng = np.random.default_rng(3)
    dirs = workloads.unit_vectors(rng, len(w["ia"])).astype(np.float32)
    dirs[::3] = w["dirs"][::3]  # a third of the pairs get a good direction
    g = ncx.query.contact(ss, poses, w["ia"], w["ib"], 0.0, init_dir=torch.as_tensor(dirs, device=ctx.device),
                          max_iter=32)
    oss = oracle_shapes(orc, w["records"], w["pool"])
    op = to_oracle_poses(w["poses"])
    r = orc.contact_support_map(oss, op, w["ia"], w["ib"], 0.0, init_dirs=dirs, max_iter=32, nthreads=8)
    capped = r["status"] == 5
    if capped.any():
        rp = orc.contact_support_map(oss, op, w["ia"][capped], w["ib"][capped], 0.0, nthreads=8)
        for key in ("has", "world1", "world2", "normal", "depth"):
            r[key][capped] = rp[key]
    assert_contacts_match(g, r, "warm-started hull pairs")


def test_warm_started_distances_match_oracle(ctx, orc):
    """Distance with init_dir and a 32-iteration cap (plain GJK recomputes capped pairs) against
    the oracle's distance run with the same directions and the same rule; also checks that the
    iteration counts (both runs summed) agree exactly."""
    w = workloads.hull_pairs(seed=14, npairs=1500)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    rng = np.random.default_rng(4)
    dirs = workloads.unit_vectors(rng, len(w["ia"])).astype(np.float32)
    dirs[::3] = w["dirs"][::3]
    dirs[1::7] = 0.0  # all-zero rows are None
    d, it = ncx.query.distance(ss, poses, w["ia"], w["ib"], init_dir=torch.as_tensor(dirs, device=ctx.device),
                               max_iter=32)
    rd, rit = orc.distance(oracle_shapes(orc, w["records"], w["pool"]), to_oracle_poses(w["poses"]), w["ia"],
                           w["ib"], nthreads=8, init_dirs=dirs, max_iter=32)
    d = d.cpu().numpy()
    assert np.array_equal(it.cpu().numpy(), rit)
    err = np.abs(d.astype(np.float64) - rd)
    assert np.all(err <= REL * np.abs(rd) + ABS), f"max err {err.max()}"
    print(f"warm-started distance: bit-exact fraction {np.mean(d == rd):.4f}")


def test_error_against_f64_reference(ctx, orc):
    """The error of the f32 GPU results against an f64 instantiation of the same algorithms on the
    same (f32-exact) inputs, reported per quantity.  The GPU equals the f32 oracle; this measures
    what f32 itself costs.  Bounds are loose sanity bounds (GJK's own stopping rule is relative
    1e-3 in f32), the printed figures are the report."""
    w = workloads.hull_pairs(seed=15, npairs=3000)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    g = ncx.query.contact(ss, poses, w["ia"], w["ib"], 0.0)
    d, _ = ncx.query.distance(ss, poses, w["ia"], w["ib"])
    o64 = oracle_shapes(orc, w["records"], w["pool"], dtype=np.float64)
    p64 = to_oracle_poses(w["poses"]).astype(np.float64)
    r = orc.contact(o64, p64, w["ia"], w["ib"], 0.0, nthreads=8)
    rd, _ = orc.distance(o64, p64, w["ia"], w["ib"], nthreads=8)
    has = g["has_contact"].cpu().numpy()
    both = has & r["has"]
    agree = np.mean(has == r["has"])
    dg = g["depth"].cpu().numpy()[both].astype(np.float64)
    de = np.abs(dg - r["depth"][both])
    ng = g["normal"].cpu().numpy()[both].astype(np.float64)
    cosang = np.clip(np.sum(ng * r["normal"][both], 1) / np.linalg.norm(ng, axis=1), -1, 1)
    ang = np.arccos(cosang)
    dd = d.cpu().numpy().astype(np.float64)
    sep = rd > 0
    rel = np.abs(dd[sep] - rd[sep]) / rd[sep]
    print(f"f32 GPU vs f64: has_contact agreement {agree:.4f}; depth |err| median {np.median(de):.2e} "
          f"p99 {np.percentile(de, 99):.2e} max {de.max():.2e}; normal angle median {np.median(ang):.2e} "
          f"p99 {np.percentile(ang, 99):.2e} rad; distance rel err median {np.median(rel):.2e} "
          f"p99 {np.percentile(rel, 99):.2e} max {rel.max():.2e}")
    assert agree > 0.99
    assert np.percentile(de, 99) < 1e-3 and np.percentile(ang, 99) < 1e-2
    assert np.percentile(rel, 99) < 1e-2


def test_raycast_grazing_rays_split_into_segments(ctx, orc):
    """Nearly horizontal rays over a flat terrain have long paths through the root box and are
    cast as several segments whose hits are reduced per ray; the result must equal the
    brute-force closest hit bit for bit (toi, normal, feature), including misses."""
    pts, tri = workloads.heightfield_mesh(seed=9, cells=64)
    pts = pts.copy()
    pts[:, 2] *= 0.05
    rng = np.random.default_rng(12)
    n = 6000
    o = np.stack([rng.uniform(-2, 8.4, n), rng.uniform(-2, 8.4, n), rng.uniform(0.02, 0.3, n)], 1).astype(np.float32)
    ang = rng.uniform(0, 2 * np.pi, n)
    dz = -rng.uniform(0.0, 0.08, n)
    d = np.stack([np.cos(ang), np.sin(ang), dz], 1)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    d[::50, 2] = 0.0  # exactly horizontal: the zero-direction slab rule
    mesh = ncx.TriMesh(pts, tri, ctx=ctx)
    toi, nrm, feat = mesh.toi_and_normal_with_ray(o, d, 30.0)
    om = orc.TriMesh(pts, tri)
    bt, bn, bf, _ = om.brute(o, d, 30.0, nthreads=8)
    assert np.array_equal(feat.cpu().numpy().view(np.uint32), bf)
    assert np.array_equal(toi.cpu().numpy(), bt)
    assert np.array_equal(nrm.cpu().numpy(), bn)
    print(f"grazing rays: {n}, hits {np.mean(bf != 0xFFFFFFFF):.3f}")
