"""The scene step (ncx_world) against the broad-phase and contact oracles over moving steps,
and hipGraph replay against eager execution."""
import numpy as np
import pytest
import torch

import ncollide_amd as ncx
from ncollide_amd import workloads
from ncollide_amd.math import to_oracle_poses
from ncollide_amd.world import World

pytestmark = pytest.mark.gpu


def oracle_shapes(orc, records, pool):
    params = np.stack([records["p0"], records["p1"], records["p2"]], 1)
    return orc.ShapeSet(records["type"], params, records["hull_offset"], records["hull_count"], pool,
                        dtype=np.float32)


def keys_of(r):
    a = r["ia"].cpu().numpy().astype(np.uint64)
    b = r["ib"].cpu().numpy().astype(np.uint64)
    return (a << np.uint64(32)) | b


@pytest.mark.parametrize("use_graph", [False, True])
def test_world_steps_match_reference(ctx, orc, use_graph):
    w = workloads.full_pipeline(seed=21, n=6000, box=40.0, nclusters=3)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    world = World(ss, margin=0.05, prediction=0.0, max_pairs=1 << 20, rebuild_every=3)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    ref_bp = orc.BroadPhase(0.05)
    oss = oracle_shapes(orc, w["records"], w["pool"])
    p_np = w["poses"].copy()
    vel = w["velocities"] * 0.2
    n_large = 0
    for step in range(7):
        world.step(poses, use_graph=use_graph)
        r = world.results()
        # reference broad phase fed the same boxes through create_proxy / deferred_set_bounding_volume
        bb = orc.shape_aabbs(oss, to_oracle_poses(p_np))
        if step == 0:
            for b in bb:
                ref_bp.create_proxy(b)
        else:
            for h in range(len(bb)):
                ref_bp.deferred_set_bounding_volume(h, bb[h])
        ref_bp.update()
        stored = world.stored_aabbs().cpu().numpy().T
        for h in range(0, len(bb), 97):
            assert np.array_equal(stored[h], ref_bp.stored_bv(h)), (step, h)
        order = np.argsort(keys_of(r))
        assert np.array_equal(keys_of(r)[order], ref_bp.pairs()), f"pairs differ at step {step}"
        ia = r["ia"].cpu().numpy()[order]
        ib = r["ib"].cpu().numpy()[order]
        ref = orc.contact(oss, to_oracle_poses(p_np), ia, ib, 0.0, nthreads=8)
        n_large += r["epa_overflow"]
        assert np.array_equal(r["has_contact"].cpu().numpy()[order], ref["has"]), f"has differs at step {step}"
        idx = np.nonzero(ref["has"])[0]
        dg = r["depth"].cpu().numpy()[order][idx].astype(np.float64)
        dr = ref["depth"][idx].astype(np.float64)
        assert np.all(np.abs(dg - dr) <= 1e-5 * np.abs(dr) + 1e-6), f"depth differs at step {step}"
        print(f"step {step}: {len(order)} pairs, {len(idx)} contacts, bit-exact {np.mean(dg == dr):.4f}, "
              f"{r['epa_overflow']} pairs past the first EPA arena")
        p_np[:, 0:3] += vel
        poses.copy_(torch.as_tensor(p_np, device=ctx.device))


def test_world_graph_replay_equals_eager(ctx):
    w = workloads.full_pipeline(seed=22, n=20000, box=60.0, nclusters=4)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    pa = torch.as_tensor(w["poses"], device=ctx.device)
    pb = pa.clone()
    wa = World(ss, margin=0.05, max_pairs=1 << 21, rebuild_every=4)
    wb = World(ss, margin=0.05, max_pairs=1 << 21, rebuild_every=4)
    vel = torch.as_tensor(w["velocities"] * 0.3, device=ctx.device)
    for step in range(9):
        wa.step(pa, use_graph=False)
        ra = wa.results()
        wb.step(pb, use_graph=True)
        rb = wb.results()
        for k in ("ia", "ib", "has_contact", "depth", "normal"):
            assert torch.equal(ra[k], rb[k]), (step, k)
        pa[:, 0:3] += vel
        pb[:, 0:3] += vel


def test_world_refit_matches_rebuild_at_4m_shapes(ctx):
    """Refitted trees (one topology kept for 5 steps) give the pair sets of trees rebuilt every
    step.  At 4M shapes the refit's cross-workgroup node list exceeds what its one-workgroup
    stage keeps in LDS, so the device-memory levels run; the smaller scenes above cover the LDS
    levels."""
    n = 1 << 22
    w = workloads.plain_pipeline(seed=31, n=n)
    rng = np.random.default_rng(31)
    poses_np = w["poses"]
    poses_np[:, 0:3] = rng.uniform(0.0, 320.0, (n, 3)).astype(np.float32)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    pa = torch.as_tensor(poses_np, device=ctx.device)
    pb = pa.clone()
    wa = World(ss, margin=0.02, max_pairs=1 << 23, rebuild_every=1)
    wb = World(ss, margin=0.02, max_pairs=1 << 23, rebuild_every=6)
    wa.set_contacts(False)
    wb.set_contacts(False)
    vel = torch.as_tensor(rng.uniform(-0.4, 0.4, (n, 3)).astype(np.float32), device=ctx.device)
    for step in range(5):
        wa.step(pa, use_graph=False)
        wb.step(pb, use_graph=False)
        ka = np.sort(keys_of(wa.results()))
        kb = np.sort(keys_of(wb.results()))
        print(f"step {step}: {len(ka)} pairs", flush=True)
        assert len(ka) > 100000
        assert np.array_equal(ka, kb), f"pairs differ at step {step}"
        pa[:, 0:3] += vel
        pb[:, 0:3] += vel


@pytest.mark.parametrize("use_graph", [False, True])
def test_world_temporal_warm_start_matches_generator(ctx, orc, use_graph):
    """The persistent narrow phase's warm start over moving steps: every pair of polyhedra starts
    GJK from its last_gjk_dir (None when the pair is new to the broad phase), as
    ConvexPolyhedronConvexPolyhedronManifoldGenerator does, and keeps its result direction for the
    next step.  The oracle's generator step is driven with a per-pair dict kept across steps; the
    scene's contacts, init directions and result directions must equal it bit for bit."""
    w = workloads.full_pipeline(seed=71, n=5000, box=32.0, nclusters=3)
    rec = w["records"]
    ss = ncx.ShapeSet(records=rec, pool=w["pool"], ctx=ctx)
    world = World(ss, margin=0.05, prediction=0.0, max_pairs=1 << 20, rebuild_every=2)
    world.set_temporal_warmstart(True)
    p_np = w["poses"].copy()
    poses = torch.as_tensor(p_np, device=ctx.device)
    oss = oracle_shapes(orc, rec, w["pool"])
    vel = w["velocities"] * 0.15
    types = rec["type"]
    last = {}  # pair key -> last_gjk_dir (the generators alive after the previous step)
    reused = 0
    for step in range(6):
        world.step(poses, use_graph=use_graph)
        r = world.results()
        init, fdir = (t.cpu().numpy() for t in world.temporal_dirs())
        ia, ib = r["ia"].cpu().numpy().astype(np.uint32), r["ib"].cpu().numpy().astype(np.uint32)
        keys = (ia.astype(np.uint64) << np.uint64(32)) | ib.astype(np.uint64)
        poly = np.isin(types[ia], (1, 3)) & np.isin(types[ib], (1, 3))
        want_init = np.zeros((len(ia), 3), np.float32)
        for k in np.nonzero(poly)[0]:
            d = last.get(int(keys[k]))
            if d is not None:
                want_init[k] = d
        assert np.array_equal(init[poly].view(np.uint32), want_init[poly].view(np.uint32)), f"step {step}: init dirs"
        reused += int(np.any(want_init[poly] != 0, 1).sum())
        op = to_oracle_poses(p_np)
        ref = orc.contact(oss, op, ia, ib, 0.0, nthreads=8)
        gen = orc.polyhedra_gjk_step(oss, op, ia[poly], ib[poly], 0.0, want_init[poly])
        for key in ("has", "world1", "world2", "normal", "depth"):
            ref[key][poly] = gen[key]
        has = r["has_contact"].cpu().numpy()
        assert np.array_equal(has, ref["has"]), f"step {step}: has_contact differs"
        for key in ("depth", "normal", "world1", "world2"):
            assert np.array_equal(r[key].cpu().numpy()[has].view(np.uint32), ref[key][has].view(np.uint32)), \
                f"step {step}: {key} differs"
        new_dir = np.where(gen["dir_set"][:, None], gen["dir"], want_init[poly])
        assert np.array_equal(fdir[poly].view(np.uint32), new_dir.view(np.uint32)), f"step {step}: result dirs"
        # the generators of this step's pairs carry over (pairs that left the broad phase drop theirs)
        last = {int(kk): d for kk, d in zip(keys[poly], new_dir) if np.any(d != 0)}
        p_np[:, :3] += vel
        poses.copy_(torch.as_tensor(p_np, device=ctx.device))
    print(f"temporal warm start: {reused} pair-steps started from last_gjk_dir over 6 steps, bit-exact")
    assert reused > 1000


@pytest.mark.parametrize("use_graph", [False, True])
def test_world_dense_cluster_pairs_match_brute_force(ctx, orc, use_graph):
    """A dense cluster: 3,000 shapes around one centre (sigma 1), so most shapes touch hundreds
    to thousands of others.  Leaves with more partners than the count pass keeps per leaf (512)
    take the write pass's re-traversal and the emit kernel's second source; the pair set of every
    step must equal brute force over the world's own stored boxes, and the first step's must
    equal the oracle's DBVTBroadPhase."""
    w = workloads.full_pipeline(seed=73, n=3000, box=40.0, nclusters=1, sigma=1.0)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    world = World(ss, margin=0.05, prediction=0.0, max_pairs=1 << 22, rebuild_every=3)
    world.set_contacts(False)
    p_np = w["poses"].copy()
    poses = torch.as_tensor(p_np, device=ctx.device)
    vel = w["velocities"] * 0.1
    oss = oracle_shapes(orc, w["records"], w["pool"])
    for step in range(5):
        world.step(poses, use_graph=use_graph)
        got = np.sort(keys_of(world.results()))
        stored = np.ascontiguousarray(world.stored_aabbs().cpu().numpy().T)
        want = np.sort(orc.pairs_brute(stored))
        assert np.array_equal(got, want), f"step {step}: {len(got)} pairs against {len(want)} by brute force"
        if step == 0:
            ref_bp = orc.BroadPhase(0.05)
            for b in orc.shape_aabbs(oss, to_oracle_poses(p_np)):
                ref_bp.create_proxy(b)
            ref_bp.update()
            assert np.array_equal(got, ref_bp.pairs())
        deg = np.bincount(np.concatenate([(got >> np.uint64(32)).astype(np.int64),
                                          (got & np.uint64(0xFFFFFFFF)).astype(np.int64)]), minlength=3000)
        over, spilled = world.spill_stats()
        print(f"step {step}: {len(got)} pairs, partners per shape max {deg.max()}, "
              f"{np.sum(deg > 512)} shapes above 512, {np.sum((deg >= 500) & (deg <= 530))} in 500..530; "
              f"leaves past 512 later partners {over}, {spilled} of them in spill chunks")
        assert deg.max() > 1500 and np.sum((deg >= 500) & (deg <= 530)) > 0
        # both ways past the 512 buffered partners run: the spill chunks (64 directories at this
        # size) and, for the leaves beyond them, the second search
        assert 0 < spilled < over
        p_np[:, :3] += vel
        poses.copy_(torch.as_tensor(p_np, device=ctx.device))


@pytest.mark.parametrize("use_graph", [False, True])
def test_world_leaf_past_the_spill_directory(ctx, orc, use_graph):
    """One cuboid overlapping 20,000 small balls, centred at the scene's lowest corner so that its
    leaf comes first in Morton order: its 20,000 later partners are more than the count pass
    keeps (512 buffered + 64 chunks of 256), so that leaf is searched a second time.  The pair set
    must equal brute force over the world's stored boxes."""
    rng = np.random.default_rng(17)
    n = 20001
    rec = np.zeros(n, workloads.SHAPE_DTYPE)
    rec["type"][0] = ncx.CUBOID
    rec["p0"][0] = rec["p1"][0] = rec["p2"][0] = 12.0
    rec["type"][1:] = ncx.BALL
    rec["p0"][1:] = rng.uniform(0.05, 0.15, n - 1)
    t = np.zeros((n, 3), np.float32)
    t[1:] = rng.uniform(1.0, 11.0, (n - 1, 3))
    q = workloads.uniform_quats(rng, n)
    q[0] = (0.0, 0.0, 0.0, 1.0)
    poses = torch.as_tensor(workloads.poses8(t, q), device=ctx.device)
    ss = ncx.ShapeSet(records=rec, pool=np.zeros((1, 3), np.float32), ctx=ctx)
    world = World(ss, margin=0.01, prediction=0.0, max_pairs=1 << 20, rebuild_every=2)
    world.set_contacts(False)
    for step in range(2):
        world.step(poses, use_graph=use_graph)
        got = np.sort(keys_of(world.results()))
        stored = np.ascontiguousarray(world.stored_aabbs().cpu().numpy().T)
        want = np.sort(orc.pairs_brute(stored))
        assert np.array_equal(got, want), f"step {step}: {len(got)} pairs against {len(want)} by brute force"
        assert np.sum((got >> np.uint64(32)) == 0) == n - 1  # the cuboid touches every ball
        over, spilled = world.spill_stats()
        assert over >= 1 and spilled < over


@pytest.mark.parametrize("n", [1, 2, 3])
def test_world_tiny_scenes(ctx, orc, n):
    """Scenes of one to three overlapping balls: the pair list and contacts of each step equal
    brute force over the stored boxes and the oracle's contacts."""
    rec = np.zeros(n, workloads.SHAPE_DTYPE)
    rec["type"][:] = ncx.BALL
    rec["p0"][:] = 0.5
    t = np.zeros((n, 3), np.float32)
    t[:, 0] = np.arange(n) * 0.6
    q = np.tile(np.array([[0, 0, 0, 1]], np.float32), (n, 1))
    p_np = workloads.poses8(t, q)
    poses = torch.as_tensor(p_np, device=ctx.device)
    ss = ncx.ShapeSet(records=rec, pool=np.zeros((1, 3), np.float32), ctx=ctx)
    world = World(ss, margin=0.05, prediction=0.0, max_pairs=64, rebuild_every=2)
    oss = oracle_shapes(orc, rec, np.zeros((1, 3), np.float32))
    for step in range(3):
        world.step(poses, use_graph=step > 0)
        r = world.results()
        got = np.sort(keys_of(r))
        stored = np.ascontiguousarray(world.stored_aabbs().cpu().numpy().T)
        assert np.array_equal(got, np.sort(orc.pairs_brute(stored)))
        assert len(got) == n - 1  # neighbours 0.6 apart with radius 0.5 touch; 1.2 apart do not
        if len(got):
            ia = r["ia"].cpu().numpy().astype(np.uint32)
            ib = r["ib"].cpu().numpy().astype(np.uint32)
            ref = orc.contact(oss, to_oracle_poses(p_np), ia, ib, 0.0)
            assert np.array_equal(r["has_contact"].cpu().numpy().astype(bool), ref["has"].astype(bool))
            assert np.array_equal(r["depth"].cpu().numpy(), np.asarray(ref["depth"], np.float32))
