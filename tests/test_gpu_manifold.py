"""GPU parity of the contact-manifold generators (ncx_contact_manifolds, through the C ABI)
against the oracle's restatement, over several updates of persistent per-interaction state:
the contacts of every manifold (count, points, normals, depths, feature ids) and every slot's
last_gjk_dir must be bit-identical.  Scenes: a clustered mix of balls, cuboids, capsules and
hulls (overlapping, random poses), and columns of shapes resting on a ground box (face-on-face
contacts with several clip points)."""
import numpy as np
import pytest
import torch

import ncollide_amd as ncx
from ncollide_amd import hull as nh
from ncollide_amd import workloads
from ncollide_amd.math import to_oracle_poses
from tests.test_gpu_world import oracle_shapes

pytestmark = pytest.mark.gpu


def hull_tris(records, pool):
    tris, off, cnt = [], np.zeros(len(records), np.uint32), np.zeros(len(records), np.uint32)
    o = 0
    for i in np.nonzero(records["type"] == ncx.CONVEX_HULL)[0]:
        h0, hc = records["hull_offset"][i], records["hull_count"][i]
        t = nh.hull_triangles(pool[h0:h0 + hc])
        tris.append(t)
        off[i], cnt[i] = o, len(t)
        o += len(t)
    return (np.concatenate(tris) if tris else np.zeros((1, 3), np.uint32)), off, cnt


def overlapping_pairs(orc, oss, poses, margin):
    bb = np.asarray(orc.shape_aabbs(oss, to_oracle_poses(poses)), np.float32).reshape(-1, 6)
    bb[:, :3] -= margin
    bb[:, 3:] += margin
    keys = orc.pairs_brute(bb)
    return (keys >> np.uint64(32)).astype(np.uint32), (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32), keys


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def run_scene(ctx, orc, w, steps, linear, ang, move, nslots=1 << 17):
    rec, pool = w["records"], w["pool"]
    ss = ncx.ShapeSet(records=rec, pool=pool, ctx=ctx)
    oss = oracle_shapes(orc, rec, pool)
    tr, toff, tcnt = hull_tris(rec, pool)
    gm = ncx.ContactManifolds(ctx, nslots)
    om = orc.ManifoldStore(nslots)
    p_np = w["poses"].copy()
    slot_of, free = {}, list(range(nslots - 1, -1, -1))
    totals = np.zeros(NCX_MAX_BINS, np.int64)
    for step in range(steps):
        ia, ib, keys = overlapping_pairs(orc, oss, p_np, 0.05)
        live = {int(k) for k in keys}
        for k in [k for k in slot_of if k not in live]:  # interactions that ended free their slot
            free.append(slot_of.pop(k))
        new = [int(k) for k in keys if int(k) not in slot_of]
        for k in new:
            slot_of[k] = free.pop()
        slots = np.array([slot_of[int(k)] for k in keys], np.uint32)
        if new:
            ns = np.array([slot_of[k] for k in new], np.uint32)
            gm.reset(ns)
            om.reset(ns)
        poses = torch.as_tensor(p_np, device=ctx.device)
        g = gm.update(ss, poses, ia, ib, slots, linear, ang, ang)
        r = om.step(oss, tr, toff, tcnt, to_oracle_poses(p_np), ia, ib, slots, linear, ang, ang)
        assert not r["overflow"]
        cnt = g["count"].cpu().numpy().astype(np.uint32)
        assert np.array_equal(cnt, r["count"]), f"step {step}: counts differ at {np.nonzero(cnt != r['count'])[0][:10]}"
        m = np.arange(ncx.manifold.MAX_CONTACTS)[None, :] < cnt[:, None]
        for key in ("world1", "world2", "normal"):
            assert np.array_equal(_bits(g[key].cpu().numpy()[m]), _bits(r[key][m])), f"step {step}: {key} differs"
        assert np.array_equal(_bits(g["depth"].cpu().numpy()[m]), _bits(r["depth"][m])), f"step {step}: depth"
        for key in ("feature1", "feature2"):
            assert np.array_equal(g[key].cpu().numpy().view(np.uint32)[m], r[key][m]), f"step {step}: {key}"
        gd = gm.dirs(torch.as_tensor(slots, device=ctx.device)).cpu().numpy()
        assert np.array_equal(_bits(gd), _bits(r["dirs"])), f"step {step}: last_gjk_dir differs"
        totals += np.bincount(np.minimum(cnt, NCX_MAX_BINS - 1), minlength=NCX_MAX_BINS)
        p_np = move(p_np, step)
    return totals


NCX_MAX_BINS = 10


def test_manifolds_clustered_mix_match_oracle(ctx, orc):
    w = workloads.full_pipeline(seed=81, n=3000, box=22.0, nclusters=2, sigma=3.0, polyhedra=True)
    vel = w["velocities"] * 0.05

    def move(p, step):
        p = p.copy()
        p[:, :3] += vel
        return p

    totals = run_scene(ctx, orc, w, 4, 0.02, 0.05, move)
    print("clustered mix: manifolds by contact count", totals.tolist())
    assert totals[1:].sum() > 3000 and totals[2:].sum() > 300


def test_manifolds_resting_stacks_match_oracle(ctx, orc):
    w = workloads.resting_stacks(seed=7, ncols=14, height=4)
    vel = w["velocities"]

    def move(p, step):
        p = p.copy()
        p[:, :3] += vel * 0.3
        return p

    totals = run_scene(ctx, orc, w, 4, 0.01, 0.1, move)
    print("resting stacks: manifolds by contact count", totals.tolist())
    assert totals[3:].sum() > 200  # face-on-face manifolds with three or more points


def test_manifold_reset_starts_fresh(ctx, orc):
    """A reset slot behaves as a new interaction: no last_gjk_dir, no tracked points."""
    w = workloads.resting_stacks(seed=9, ncols=6, height=3)
    rec, pool = w["records"], w["pool"]
    ss = ncx.ShapeSet(records=rec, pool=pool, ctx=ctx)
    oss = oracle_shapes(orc, rec, pool)
    ia, ib, _ = overlapping_pairs(orc, oss, w["poses"], 0.05)
    n = len(ia)
    gm = ncx.ContactManifolds(ctx, 2 * n)
    slots = np.arange(n, dtype=np.uint32)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    gm.reset(np.arange(2 * n, dtype=np.uint32))
    gm.update(ss, poses, ia, ib, slots, 0.01)
    gm.reset(slots)
    a = gm.update(ss, poses, ia, ib, slots, 0.01)
    b = gm.update(ss, poses, ia, ib, slots + n, 0.01)  # never-used slots
    for key in ("count", "depth", "feature1", "feature2"):
        assert torch.equal(a[key], b[key]), key
