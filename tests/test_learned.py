"""The learned warm start on the coherent hull-pair workload and in the scene step.

The network is trained on the GPU by the library's own kernels on plain GJK/EPA results, then used
as init_dir for contact (prediction 0) and distance, with GJK capped at 32 iterations and plain GJK
on the cap.  Checked here:
  * parity: the kernels driven by the network's own predicted directions equal the oracle given
    the same directions (bit for bit), on the hull-pair distribution of the learned workload;
  * accuracy: against the f64 oracle (the error reference), the learned results are as accurate
    as plain GJK's (GJK's termination tests are unchanged; different start directions end at
    different points inside the same relative tolerance);
  * gain: GJK iterations per distance query drop;
  * the scene step's learned leg (World.set_warmstart): contacts equal the oracle given the
    directions the step used, eager and graph-replayed.
"""
import numpy as np
import pytest
import torch

import ncollide_amd as ncx
from ncollide_amd import learned, query, workloads
from ncollide_amd.math import to_oracle_poses
from ncollide_amd.warmstart import descriptors
from ncollide_amd.world import World
from tests.test_gpu_parity import oracle_shapes

pytestmark = pytest.mark.gpu

CAP = 32


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def trained(ctx):
    feat, y = learned.training_set(ctx, seed=51, nsamples=160_000, pool_hulls=8192)
    mlp, hist, _ = learned.train(ctx, feat, y, epochs=6)
    return mlp, hist


def _oracle_capped_contacts(orc, oss, op, ia, ib, dirs, prediction=0.0):
    """contact_support_map_support_map_with_params from `dirs` with GJK capped at CAP, plain GJK
    (init_dir = None) recomputing the capped pairs: the capped kernels' rule."""
    r = orc.contact_support_map(oss, op, ia, ib, prediction, init_dirs=dirs, max_iter=CAP, nthreads=8)
    capped = r["status"] == 5
    if capped.any():
        rp = orc.contact_support_map(oss, op, ia[capped], ib[capped], prediction, nthreads=8)
        for key in ("has", "world1", "world2", "normal", "depth"):
            r[key][capped] = rp[key]
    return r, capped


def test_learned_training_converges(trained):
    mlp, hist = trained
    assert hist[-1] < 0.5 * hist[0]


def test_learned_dirs_through_capped_kernels_match_oracle(ctx, orc, trained):
    mlp, _ = trained
    w = workloads.hull_pairs(seed=61, npairs=6000)  # the learned workload's hull distribution
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    desc = descriptors(ss)
    dirs = mlp.predict_dirs(ss, poses, desc, w["ia"], w["ib"])
    g = query.contact(ss, poses, w["ia"], w["ib"], 0.0, init_dir=dirs, max_iter=CAP)
    gd, git = query.distance(ss, poses, w["ia"], w["ib"], init_dir=dirs, max_iter=CAP)
    dh = dirs.cpu().numpy()
    assert np.all(np.abs(np.linalg.norm(dh, axis=1) - 1) < 1e-5)
    oss = oracle_shapes(orc, w["records"], w["pool"])
    op = to_oracle_poses(w["poses"])
    r, capped = _oracle_capped_contacts(orc, oss, op, w["ia"], w["ib"], dh)
    has = g["has_contact"].cpu().numpy()
    assert np.array_equal(has, r["has"])
    for key, gk in (("depth", "depth"), ("normal", "normal"), ("world1", "world1"), ("world2", "world2")):
        assert np.array_equal(_bits(g[gk].cpu().numpy()[has]), _bits(r[key][has])), f"contact {key} differs"
    rd, rit = orc.distance(oss, op, w["ia"], w["ib"], nthreads=8, init_dirs=dh, max_iter=CAP)
    assert np.array_equal(_bits(gd.cpu().numpy()), _bits(rd)), "distance differs"
    assert np.array_equal(git.cpu().numpy(), rit), "distance GJK iterations differ"
    print(f"learned dirs through the capped kernels: {len(has)} pairs, {has.sum()} contacts, "
          f"{capped.sum()} contact runs capped, bit-exact")


def test_learned_accuracy_matches_plain_against_f64(ctx, orc, trained):
    mlp, _ = trained
    w = workloads.hull_pairs(seed=62, npairs=8192)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    desc = descriptors(ss)
    rng = np.random.default_rng(63)
    v, om = workloads.coherent_motion(rng, len(w["records"]))
    poses_np = w["poses"]
    oss64 = oracle_shapes(orc, w["records"], w["pool"], dtype=np.float64)
    errs = {"plain": ([], []), "learned": ([], [])}
    for step in range(3):
        poses = torch.as_tensor(poses_np, device=ctx.device)
        dirs = mlp.predict_dirs(ss, poses, desc, w["ia"], w["ib"])
        out = {"plain": (query.contact(ss, poses, w["ia"], w["ib"], 0.0), query.distance(ss, poses, w["ia"], w["ib"])),
               "learned": (query.contact(ss, poses, w["ia"], w["ib"], 0.0, init_dir=dirs, max_iter=CAP),
                           query.distance(ss, poses, w["ia"], w["ib"], init_dir=dirs, max_iter=CAP))}
        op = to_oracle_poses(poses_np).astype(np.float64)
        t_d, _ = orc.distance(oss64, op, w["ia"], w["ib"], nthreads=8)
        t_c = orc.contact(oss64, op, w["ia"], w["ib"], 0.0, nthreads=8)
        for key, (c, (d, _)) in out.items():
            d = d.cpu().numpy().astype(np.float64)
            sep = t_d > 0
            errs[key][0].append(np.abs(d[sep] - t_d[sep]) / np.maximum(t_d[sep], 1e-3))
            h = c["has_contact"].cpu().numpy() & t_c["has"]
            dep = c["depth"].cpu().numpy().astype(np.float64)[h]
            errs[key][1].append(np.abs(dep - t_c["depth"][h]) / np.maximum(np.abs(t_c["depth"][h]), 1e-3))
        poses_np = workloads.step_poses(poses_np, v, om)
    p99 = {k: (np.percentile(np.concatenate(e[0]), 99), np.percentile(np.concatenate(e[1]), 99)) for k, e in errs.items()}
    med = {k: (np.median(np.concatenate(e[0])), np.median(np.concatenate(e[1]))) for k, e in errs.items()}
    print(f"relative error vs f64, distance p99 plain {p99['plain'][0]:.3g} learned {p99['learned'][0]:.3g} "
          f"(median {med['plain'][0]:.3g} / {med['learned'][0]:.3g}); depth p99 plain {p99['plain'][1]:.3g} "
          f"learned {p99['learned'][1]:.3g}")
    # as accurate as plain GJK: within 25% of plain's p99 (both end inside GJK's relative tolerance)
    assert p99["learned"][0] <= 1.25 * p99["plain"][0] + 1e-7
    assert p99["learned"][1] <= 1.25 * p99["plain"][1] + 1e-7


def test_learned_warm_start_gains_over_plain(ctx, trained):
    mlp, _ = trained
    res = learned.evaluate_coherent(ctx, mlp, seed=52, npairs=8192, steps=4, placer_iters=16)
    p, l, o, t = res["plain"], res["learned"], res["oracle_dir"], res["temporal"]
    print(f"distance GJK it/pair: plain {p['distance_gjk_iters_per_pair']:.2f}, learned "
          f"{l['distance_gjk_iters_per_pair']:.2f}, temporal {t['distance_gjk_iters_per_pair']:.2f}, exact "
          f"direction {o['distance_gjk_iters_per_pair']:.2f}; contact GJK it/pair: plain "
          f"{p['contact_gjk_iters_per_pair']:.2f}, learned {l['contact_gjk_iters_per_pair']:.2f}")
    assert l["distance_gjk_iters_per_pair"] < 0.95 * p["distance_gjk_iters_per_pair"]


@pytest.mark.parametrize("use_graph", [False, True])
def test_world_learned_leg_matches_oracle(ctx, orc, trained, use_graph):
    """The scene step with the learned warm start: polytope pairs start from the network's
    direction (capped at 32 iterations, plain GJK on the cap), every other pair as plain."""
    mlp, _ = trained
    w = workloads.full_pipeline(seed=64, n=5000, box=35.0, nclusters=3)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    world = World(ss, margin=0.05, prediction=0.0, max_pairs=1 << 20, rebuild_every=2)
    world.set_warmstart(mlp)
    p_np = w["poses"].copy()
    poses = torch.as_tensor(p_np, device=ctx.device)
    oss = oracle_shapes(orc, w["records"], w["pool"])
    vel = w["velocities"] * 0.2
    types = w["records"]["type"]
    warm_total = 0
    for step in range(4):
        world.step(poses, use_graph=use_graph)
        r = world.results()
        dirs = world.warmstart_dirs().cpu().numpy()
        ia, ib = r["ia"].cpu().numpy().astype(np.uint32), r["ib"].cpu().numpy().astype(np.uint32)
        op = to_oracle_poses(p_np)
        warm = np.any(dirs != 0, 1)
        poly = np.isin(types[ia], (1, 3)) & np.isin(types[ib], (1, 3))
        assert np.array_equal(warm, poly), "directions are set on exactly the polytope pairs"
        ref = orc.contact(oss, op, ia, ib, 0.0, nthreads=8)
        if warm.any():
            rw, _ = _oracle_capped_contacts(orc, oss, op, ia[warm], ib[warm], dirs[warm])
            for key in ("has", "world1", "world2", "normal", "depth"):
                ref[key][warm] = rw[key]
        has = r["has_contact"].cpu().numpy()
        assert np.array_equal(has, ref["has"]), f"step {step}: has_contact differs"
        assert np.array_equal(_bits(r["depth"].cpu().numpy()[has]), _bits(ref["depth"][has])), f"step {step}: depth"
        assert np.array_equal(_bits(r["normal"].cpu().numpy()[has]), _bits(ref["normal"][has])), f"step {step}: normal"
        warm_total += warm.sum()
        p_np[:, :3] += vel
        poses.copy_(torch.as_tensor(p_np, device=ctx.device))
    print(f"scene learned leg: {warm_total} warm-started polytope pairs over 4 steps, bit-exact")
    assert warm_total > 1000
