"""GPU parity of the broad phase's remove handler, collision-group pair filter and batched
interference queries (boxes, points, rays), and of the world-level ray queries
(first_interference_with_ray, interferences_with_ray), all through the C ABI, against the
oracle's DBVTBroadPhase and brute force over the same stored boxes."""
import numpy as np
import pytest
import torch

import ncollide_amd as ncx
from ncollide_amd import workloads
from ncollide_amd.math import to_oracle_poses
from ncollide_amd.world import World
from tests.test_gpu_parity import oracle_shapes

pytestmark = pytest.mark.gpu


def _boxes(c, r):
    return np.concatenate([c - r, c + r], 1).astype(np.float32)


def _csr_sets(offs, handles):
    offs = offs.cpu().numpy().astype(np.int64)
    hs = handles.cpu().numpy().view(np.uint32)
    return [np.sort(hs[offs[k]:offs[k + 1]]) for k in range(len(offs) - 1)]


def test_broad_phase_remove_and_recreate_events(ctx, orc):
    """remove() reports the pairs it drops (the reference's handler) and they are not reported
    again; handles are reused LIFO by create_proxy and the new pairs start at the next update."""
    rng = np.random.default_rng(31)
    n = 2500
    c = rng.uniform(0, 25, (n, 3)).astype(np.float32)
    r = rng.uniform(0.3, 0.9, (n, 1)).astype(np.float32)
    bp = ncx.DBVTBroadPhase(0.05, ctx=ctx)
    ref = orc.BroadPhase(0.05)
    hs = list(bp.create_proxies(_boxes(c, r)))
    assert hs == [ref.create_proxy(b) for b in _boxes(c, r)]
    alive = set(hs)
    for step in range(5):
        bp.update()
        ref.update()
        assert np.array_equal(bp.pairs(), ref.pairs()), f"pairs differ at step {step}"
        s, t = bp.events()
        rs, rt = ref.events()
        assert np.array_equal(s, rs), f"started differ at step {step}"
        assert np.array_equal(np.sort(np.concatenate([t, removed])) if step else t, rt), f"stopped at step {step}"
        # remove some proxies, then create new ones (they take the freed handles, last freed first)
        gone = rng.choice(sorted(alive), 40, replace=False).astype(np.uint32)
        removed = bp.remove(gone)
        ref.remove(gone)
        alive -= set(gone.tolist())
        cn = c[gone] + rng.normal(0, 0.3, (len(gone), 3)).astype(np.float32)  # near the old places
        nb = _boxes(cn, r[gone])
        new = list(bp.create_proxies(nb))
        assert new == [ref.create_proxy(b) for b in nb]
        assert set(new) == set(gone.tolist()), "freed handles are reused"
        alive |= set(new)
        c[new] = cn
        # move everything that is alive
        c = c + rng.normal(0, 0.08, c.shape).astype(np.float32)
        al = np.array(sorted(alive), np.uint32)
        bp.deferred_set_bounding_volumes(al, _boxes(c[al], r[al]))
        for h in al:
            ref.deferred_set_bounding_volume(int(h), _boxes(c[h:h + 1], r[h:h + 1])[0])
    print(f"remove handler: last step {len(removed)} pairs dropped by remove()")


def test_broad_phase_collision_groups_filter(ctx, orc):
    rng = np.random.default_rng(32)
    n = 3000
    c = rng.uniform(0, 25, (n, 3)).astype(np.float32)
    r = rng.uniform(0.3, 0.9, (n, 1)).astype(np.float32)
    bp = ncx.DBVTBroadPhase(0.05, ctx=ctx)
    ref = orc.BroadPhase(0.05)
    hs = bp.create_proxies(_boxes(c, r))
    for b in _boxes(c, r):
        ref.create_proxy(b)
    # four groups: membership one group each; whitelist 2-3 groups; a few blacklists
    grp = rng.integers(0, 4, n)
    mem = (1 << grp).astype(np.uint32)
    wl = np.where(rng.random(n) < 0.5, 0xF, 0b0101 << (grp % 2)).astype(np.uint32)
    bl = np.where(rng.random(n) < 0.1, 1 << ((grp + 1) % 4), 0).astype(np.uint32)
    bp.set_collision_groups(hs, mem, wl, bl)
    for h in range(n):
        ref.set_collision_groups(h, mem[h], wl[h], bl[h])
    for step in range(4):
        bp.update()
        ref.update()
        assert np.array_equal(bp.pairs(), ref.pairs()), f"pairs differ at step {step}"
        s, t = bp.events()
        rs, rt = ref.events()
        assert np.array_equal(s, rs) and np.array_equal(t, rt), f"events differ at step {step}"
        c = c + rng.normal(0, 0.08, c.shape).astype(np.float32)
        bp.deferred_set_bounding_volumes(hs, _boxes(c, r))
        for h in range(n):
            ref.deferred_set_bounding_volume(h, _boxes(c[h:h + 1], r[h:h + 1])[0])
    allp = orc.pairs_brute(_boxes(c, r))
    print(f"groups: {len(bp.pairs())} pairs kept")


def test_broad_phase_interference_queries(ctx, orc):
    rng = np.random.default_rng(33)
    n = 4000
    c = rng.uniform(0, 30, (n, 3)).astype(np.float32)
    r = rng.uniform(0.2, 1.0, (n, 1)).astype(np.float32)
    bp = ncx.DBVTBroadPhase(0.05, ctx=ctx)
    hs = bp.create_proxies(_boxes(c, r))
    bp.update()
    # a few removals after the update: removed proxies leave the queries at once (the reference
    # takes their leaves out of its trees inside remove)
    gone = np.array([3, 17, 400, 2500], np.uint32)
    bp.remove(gone)
    stored = np.stack([bp.proxy(h) if h not in gone else np.full(6, np.nan, np.float32) for h in range(n)])
    live = ~np.isnan(stored[:, 0])
    nq = 3000
    qc = rng.uniform(-2, 32, (nq, 3)).astype(np.float32)
    qb = _boxes(qc, rng.uniform(0.0, 1.5, (nq, 1)).astype(np.float32))
    qb[::10, 3:] = qb[::10, :3]  # degenerate (point-like) boxes
    got = _csr_sets(*bp.interferences_with_bounding_volumes(qb))
    for k in range(nq):
        m = live & np.all(stored[:, :3] <= qb[k, 3:], 1) & np.all(stored[:, 3:] >= qb[k, :3], 1)
        assert np.array_equal(got[k], np.nonzero(m)[0].astype(np.uint32)), f"box query {k}"
    pts = np.concatenate([qc, stored[live][:500, :3]])  # include points exactly on box corners
    got = _csr_sets(*bp.interferences_with_points(pts))
    for k in range(len(pts)):
        m = live & np.all(stored[:, :3] <= pts[k], 1) & np.all(stored[:, 3:] >= pts[k], 1)
        assert np.array_equal(got[k], np.nonzero(m)[0].astype(np.uint32)), f"point query {k}"
    ro = rng.uniform(-5, 35, (nq, 3)).astype(np.float32)
    rd = rng.normal(size=(nq, 3)).astype(np.float32)
    rd[::7, 0] = 0.0  # zero direction components
    got = _csr_sets(*bp.interferences_with_rays(ro, rd, 20.0))
    bx = np.where(live[:, None], stored, np.array([1e30] * 6, np.float32))  # removed: a point out of reach
    dummy = orc.ShapeSet(np.zeros(n, np.uint32), np.ones((n, 3)), np.zeros(n), np.zeros(n), np.zeros((0, 3)),
                         dtype=np.float32)
    cnt, h, _, _ = orc.world_rays(dummy, np.tile([0, 0, 0, 0, 0, 0, 1.0], (n, 1)), bx, ro, rd, np.float32(20.0), 2)
    offs = np.concatenate([[0], np.cumsum(cnt)])
    hits = 0
    for k in range(nq):
        assert np.array_equal(got[k], h[offs[k]:offs[k + 1]]), f"ray query {k}"
        hits += len(got[k])
    print(f"bp queries: {nq} boxes, {len(pts)} points, {nq} rays ({hits} ray hits)")


@pytest.mark.parametrize("use_graph", [False, True])
def test_world_ray_queries_match_oracle(ctx, orc, use_graph):
    w = workloads.full_pipeline(seed=34, n=6000, box=30.0, nclusters=3)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    world = World(ss, margin=0.05, prediction=0.0, max_pairs=1 << 20, rebuild_every=2)
    p_np = w["poses"].copy()
    poses = torch.as_tensor(p_np, device=ctx.device)
    vel = w["velocities"] * 0.3
    rng = np.random.default_rng(35)
    oss = oracle_shapes(orc, w["records"], w["pool"])
    for step in range(3):
        world.step(poses, use_graph=use_graph)
        world.sync()
        stored = world.stored_aabbs().cpu().numpy().T.copy()
        nr = 4000
        tgt = p_np[rng.integers(0, len(p_np), nr), :3]
        ro = (tgt + rng.normal(0, 6, (nr, 3))).astype(np.float32)
        rd = ((tgt - ro) * rng.uniform(0.2, 1.5, (nr, 1))).astype(np.float32)
        rd[::11] = rng.normal(size=(len(rd[::11]), 3)).astype(np.float32)
        op = to_oracle_poses(p_np)
        gh, gt, gn = world.first_interference_with_rays(poses, ro, rd, 40.0)
        gh = gh.cpu().numpy().view(np.uint32)
        rh, rt, rn = orc.world_rays(oss, op, stored, ro, rd, np.float32(40.0), 0)
        assert np.array_equal(gh, rh), f"step {step}: first-hit handles differ at {np.nonzero(gh != rh)[0][:10]}"
        assert np.array_equal(gt.cpu().numpy().view(np.uint32), rt.view(np.uint32)), f"step {step}: toi"
        assert np.array_equal(gn.cpu().numpy().view(np.uint32), rn.view(np.uint32)), f"step {step}: normals"
        for aabb_only, mode in ((False, 1), (True, 2)):
            offs, hs, ts, ns = world.interferences_with_rays(poses, ro, rd, 40.0, aabb_only=aabb_only)
            cnt, h, t, n_ = orc.world_rays(oss, op, stored, ro, rd, np.float32(40.0), mode)
            o = offs.cpu().numpy().astype(np.int64)
            assert np.array_equal(np.diff(o), cnt), f"step {step}: counts differ (aabb_only={aabb_only})"
            ro_ = np.concatenate([[0], np.cumsum(cnt)])
            hs_, ts_ = hs.cpu().numpy().view(np.uint32), ts.cpu().numpy()
            for k in range(0, nr, 7):
                sl = slice(o[k], o[k + 1])
                perm = np.argsort(hs_[sl])
                assert np.array_equal(hs_[sl][perm], h[ro_[k]:ro_[k + 1]])
                if not aabb_only:
                    assert np.array_equal(ts_[sl][perm].view(np.uint32), t[ro_[k]:ro_[k + 1]].view(np.uint32))
        qb = _boxes(tgt[:1000].astype(np.float32), np.full((1000, 1), 0.7, np.float32))
        got = _csr_sets(*world.interferences_with_aabbs(qb))
        for k in range(0, 1000, 5):
            m = np.all(stored[:, :3] <= qb[k, 3:], 1) & np.all(stored[:, 3:] >= qb[k, :3], 1)
            assert np.array_equal(got[k], np.nonzero(m)[0].astype(np.uint32))
        print(f"step {step}: first hits {np.mean(rh != 0xFFFFFFFF):.3f}, ray interferences {len(h)}")
        p_np[:, :3] += vel
        poses.copy_(torch.as_tensor(p_np, device=ctx.device))


def test_first_interference_with_ray_reference_case(ctx):
    """tests/geometry/first_interference_with_ray.rs:11-76: two unit balls, the ray misses the first
    and hits the second; one interference, and the first one is that ball."""
    rec = np.zeros(2, ncx.shape.SHAPE_DTYPE)
    rec["type"] = 0
    rec["p0"] = 1.0
    s, cpi = np.sin(np.pi / 2), np.cos(np.pi / 2)  # rotation by pi about y (scaled axis y * pi)
    poses = np.array([[1, 1, 0, 0, 0, s, 0, cpi], [10, 11.8, 0, 0, 0, s, 0, cpi]], np.float32)
    ss = ncx.ShapeSet(records=rec, pool=np.zeros((0, 3), np.float32), ctx=ctx)
    world = World(ss, margin=0.01, prediction=0.0, max_pairs=16)
    pt = torch.as_tensor(poses, device=ctx.device)
    world.step(pt, use_graph=False)
    world.sync()
    o = np.array([[0.0, 1.8, 0.0]], np.float32)
    d = (np.array([[1.0, 1.0, 0.0]]) / np.sqrt(2.0)).astype(np.float32)
    offs, hs, ts, ns = world.interferences_with_rays(pt, o, d, 3.4e38)
    assert offs.cpu().numpy().tolist() == [0, 1] and hs.cpu().numpy().tolist() == [1]
    h, t, n = world.first_interference_with_rays(pt, o, d, 3.4e38)
    # toi of the ray on ball 2 (closed form)
    dc = o[0].astype(np.float64) - poses[1, :3]
    b = dc @ d[0]
    toi_ball = -b - np.sqrt(b * b - (dc @ dc - 1.0))
    assert h.cpu().numpy().tolist() == [1] and abs(t.cpu().numpy()[0] - toi_ball) < 1e-4


@pytest.mark.parametrize("nr", [1023, 1024, 3000])
def test_world_first_hits_either_side_of_the_sorted_batch_size(ctx, orc, nr):
    """Batches of 1,024 rays or more are taken in origin/octant key order, smaller ones in input
    order: both give the oracle's first hits in each ray's own slot."""
    w = workloads.full_pipeline(seed=41, n=3000, box=20.0, nclusters=2)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    world = World(ss, margin=0.05, prediction=0.0, max_pairs=1 << 18)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    world.step(poses, use_graph=False)
    world.sync()
    stored = world.stored_aabbs().cpu().numpy().T.copy()
    rng = np.random.default_rng(nr)
    tgt = w["poses"][rng.integers(0, len(w["poses"]), nr), :3]
    ro = (tgt + rng.normal(0, 5, (nr, 3))).astype(np.float32)
    rd = ((tgt - ro) * rng.uniform(0.2, 1.5, (nr, 1))).astype(np.float32)
    rd[::5] = -rd[::5]  # every octant, and rays that leave the scene
    gh, gt, gn = world.first_interference_with_rays(poses, ro, rd, 40.0)
    oss = oracle_shapes(orc, w["records"], w["pool"])
    rh, rt, rn = orc.world_rays(oss, to_oracle_poses(w["poses"]), stored, ro, rd, np.float32(40.0), 0)
    assert np.array_equal(gh.cpu().numpy().view(np.uint32), rh)
    assert np.array_equal(gt.cpu().numpy().view(np.uint32), rt.view(np.uint32))
    assert np.array_equal(gn.cpu().numpy().view(np.uint32), rn.view(np.uint32))
    assert np.mean(rh != 0xFFFFFFFF) > 0.3
