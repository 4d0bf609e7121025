"""Benchmark of the collision hot path on one MI355X.

Headline line: TriMesh ray casting, 1,048,576 rays against a 257x257 heightfield
TriMesh (131,072 triangles), max_toi 100, first-hit toi + normal + triangle id.
One step = one cast of the whole ray batch; rays and mesh are resident in HBM.

Also reported in the same JSON line ("secondary"): batched contact + distance on
262,144 ConvexHull pairs (with its CPU baseline), the learned warm start against
plain GJK over 64 coherent steps, the 1,048,576-shape scene step, and the
1,024-shape plain pipeline step.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
import argparse
import csv
import json
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time

import ctypes as C

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK = 157.3e12  # dense FP32 vector peak of an MI355X, FLOP/s
HBM_PEAK = 8.0e12  # bytes/s, MI355X HBM3E


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--quick", action="store_true", help="skip the secondary workloads")
    p.add_argument("--only", choices=["hulls", "learned", "scene", "plain", "queries", "manifolds"], default=None,
                   help="run one secondary workload alone (after the ray-cast headline); for profiling")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--copy-calibration", action="store_true",
                   help="copy 256 MiB device to device once before the timed casts (the reference dispatch "
                        "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE passes over the ray cast")
    p.add_argument("--rays", type=int, default=1 << 20)
    p.add_argument("--hull-pairs", type=int, default=1 << 18)
    p.add_argument("--train-samples", type=int, default=2_000_000)
    p.add_argument("--c4-steps", type=int, default=64)
    p.add_argument("--scene-shapes", type=int, default=1 << 20)
    p.add_argument("--scene-steps", type=int, default=32)
    p.add_argument("--rebuild-every", type=int, default=2,
                   help="scene step: LBVH rebuilt every k steps, refit in between")
    p.add_argument("--scene-max-pairs", type=int, default=40_000_000)
    return p.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def bench_rays(args, ncx, torch, ctx):
    from ncollide_amd import workloads
    pts, tri = workloads.heightfield_mesh(seed=2)
    o, d = workloads.heightfield_rays(seed=2, n=args.rays)
    mesh = ncx.TriMesh(pts, tri, ctx=ctx)
    dev = ctx.device
    soa = [torch.as_tensor(np.ascontiguousarray(a), device=dev) for a in (o[:, 0], o[:, 1], o[:, 2], d[:, 0],
                                                                        d[:, 1], d[:, 2])]
    n = args.rays
    toi = torch.empty(n, dtype=torch.float32, device=dev)
    nrm = torch.empty((n, 3), dtype=torch.float32, device=dev)
    feat = torch.empty(n, dtype=torch.int32, device=dev)
    pose = mesh._pose(None)
    lib = ctx.lib
    from ncollide_amd import _lib
    rays_p = _lib.ptr6(soa)
    out = (_lib.ptr(toi), _lib.ptr(nrm), _lib.ptr(feat))
    pose_p = _lib.ptr(pose)

    def launch():
        _lib.check(lib.ncx_trimesh_cast_rays_async(mesh.h, pose_p, rays_p, n, 100.0, *out))

    # traversal counts for the algorithmic-bytes model (same traversal, counting pass)
    nodes, tris = mesh.cast_stats(None, None, 100.0, soa=soa)
    nodes = nodes.to(torch.int64).sum().item()
    tris = tris.to(torch.int64).sum().item()
    # bytes: ray 24 + hit 20 per ray, 32 per child box tested (2 per internal node visited), 36 per triangle tested
    alg_bytes = 44 * n + 64 * nodes + 36 * tris
    for _ in range(args.warmup):
        launch()
    torch.cuda.synchronize()
    # whole casts, back to back (the value): barrier + synchronize on both sides
    t0 = time.perf_counter()
    for _ in range(args.steps):
        launch()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _lib.check(lib.ncx_trimesh_check(mesh.h))
    wall_ms = (t1 - t0) * 1e3 / args.steps
    # the traversal kernel alone, from HIP events the library records around it on its stream
    _lib.check(lib.ncx_trimesh_set_timing(mesh.h, 1))
    kms = []
    for _ in range(args.steps):
        launch()
        v = C.c_float()
        _lib.check(lib.ncx_trimesh_kernel_ms(mesh.h, C.byref(v)))
        kms.append(v.value)
    _lib.check(lib.ncx_trimesh_set_timing(mesh.h, 0))
    torch.cuda.synchronize()
    avg_kernel_s = float(np.mean(kms)) / 1e3
    hits = float((feat != -1).float().mean().item())
    ns = min(n, 65536)
    return dict(n=n, wall_ms=wall_ms, kernel_ms=float(np.mean(kms)), kernel_ms_min=float(np.min(kms)),
                alg_bytes=alg_bytes, nodes_per_ray=nodes / n, tris_per_ray=tris / n,
                achieved=alg_bytes / avg_kernel_s, hits=hits, pts=pts, tri=tri, o=o, d=d,
                sample=(toi[:ns].cpu().numpy(), nrm[:ns].cpu().numpy(), feat[:ns].cpu().numpy().view(np.uint32)))


def cpu_accuracy_rays(r):
    """Error of the GPU's f32 hits against the reference algorithm in f64 (the oracle's BVT
    best-first search with Real = double) on the first 65,536 rays: relative toi error, normal
    angle, and how often the hit triangle and the hit/miss decision agree."""
    from oracle import pyoracle as orc
    toi, nrm, feat = r["sample"]
    ns = len(toi)
    m = orc.TriMesh(r["pts"].astype(np.float64), r["tri"], dtype=np.float64)
    t64, n64, f64, _ = m.cast(r["o"][:ns].astype(np.float64), r["d"][:ns].astype(np.float64), 100.0,
                              nthreads=min(os.cpu_count() or 1, 64))
    hit = (feat != 0xFFFFFFFF) & (f64 != 0xFFFFFFFF)
    rel = np.abs(toi[hit].astype(np.float64) - t64[hit]) / np.maximum(np.abs(t64[hit]), 1e-12)
    cosang = np.clip(np.sum(nrm[hit].astype(np.float64) * n64[hit], 1), -1.0, 1.0)
    ang = np.arccos(cosang)
    return {"rays": int(ns), "hit_agreement": float(np.mean((feat != 0xFFFFFFFF) == (f64 != 0xFFFFFFFF))),
            "triangle_agreement": float(np.mean(feat[hit] == f64[hit])),
            "toi_rel_err_p50": float(np.percentile(rel, 50)), "toi_rel_err_p99": float(np.percentile(rel, 99)),
            "toi_rel_err_max": float(rel.max()), "normal_angle_p99_rad": float(np.percentile(ang, 99))}


def cpu_baseline_rays(pts, tri, o, d):
    """The oracle's restatement of the reference path (median BVT + best-first search per ray),
    on host cores: all hardware threads on 2^18 rays, and 1 thread on 2^15 rays."""
    from oracle import pyoracle as orc
    m = orc.TriMesh(pts, tri)
    threads = min(os.cpu_count() or 1, 64)
    ns = min(len(o), 1 << 18)
    t = time.perf_counter()
    m.cast(o[:ns], d[:ns], 100.0, nthreads=threads)
    mt = ns / (time.perf_counter() - t)
    n1 = min(len(o), 1 << 15)
    t = time.perf_counter()
    m.cast(o[:n1], d[:n1], 100.0, nthreads=1)
    st = n1 / (time.perf_counter() - t)
    return dict(value=mt, unit="rays/s", cores=threads, kind="port",
                sample=f"first {ns} of the same rays, reference algorithm (median-split BVT + best-first search), "
                       f"{threads} threads", single_thread=st)


def gpu_placer(ncx, torch, ctx):
    """Signed distance of hull pairs for workload placement: -depth of query::contact with a large
    prediction (pairs farther apart than it count as 10)."""
    state = {}

    def placer(records, pool, ia, ib):
        dev = ctx.device
        state["ss"] = ncx.ShapeSet(records=records, pool=pool, ctx=ctx)
        state["ia"] = torch.as_tensor(ia.astype(np.int32), device=dev)
        state["ib"] = torch.as_tensor(ib.astype(np.int32), device=dev)

        def signed_distance(poses8):
            g = ncx.query.contact(state["ss"], torch.as_tensor(poses8, device=dev), state["ia"], state["ib"], 10.0)
            sd = (-g["depth"]).cpu().numpy()
            sd[~g["has_contact"].cpu().numpy()] = 10.0
            return sd
        return signed_distance
    return placer, state


def _timed(torch, fn, reps):
    """Median wall time of fn() (each call completes on the GPU before the next starts)."""
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def bench_hulls(args, ncx, torch, ctx):
    """262,144 ConvexHull pairs (N~U{8..256} points per hull, half penetrating by U[0,0.1], half
    apart by U[0,0.5]): query::contact at prediction 0 and query::distance per pair, one batched
    call each.  Roofline: compulsory HBM bytes per pair and call (two poses and two shape records,
    32 B each, the pair's indices, both hulls' vertices once, the outputs)."""
    from ncollide_amd import workloads
    t = time.perf_counter()
    placer, state = gpu_placer(ncx, torch, ctx)
    w = workloads.hull_pairs(seed=3, npairs=args.hull_pairs, placer=placer, bisect_iters=20)
    gen_s = time.perf_counter() - t
    ss, ia, ib = state["ss"], state["ia"], state["ib"]
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    reps = max(3, args.steps // 4)
    tc = _timed(torch, lambda: ncx.query.contact(ss, poses, ia, ib, 0.0), reps)
    td = _timed(torch, lambda: ncx.query.distance(ss, poses, ia, ib), reps)
    g = ncx.query.contact(ss, poses, ia, ib, 0.0)
    d, dit = ncx.query.distance(ss, poses, ia, ib)
    n = len(w["ia"])
    nverts = w["records"]["hull_count"].astype(np.float64)
    verts_per_pair = float(nverts[w["ia"]].mean() + nverts[w["ib"]].mean())
    per_pair = 2 * (64 + 64 + 8) + 2 * 12 * verts_per_pair + (1 + 36 + 4 + 4) + (4 + 4)
    achieved = per_pair * n / (tc + td)
    cit = g["gjk_iters"].float().mean().item()
    dits = dit.float().mean().item()
    return dict(pairs=n, contact_ms=tc * 1e3, distance_ms=td * 1e3, pairs_per_s=n / (tc + td),
                first_tier_variant_by_bucket={"contact": ss.tuning("contact")[0], "distance": ss.tuning("distance")[0]},
                contact_pairs_per_s=n / tc, distance_pairs_per_s=n / td,
                contact_fraction=g["has_contact"].float().mean().item(),
                contact_gjk_iters=cit, distance_gjk_iters=dits, hull_vertices_per_pair=verts_per_pair,
                roofline={"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                          "frac": achieved / HBM_PEAK, "bytes_per_pair": per_pair},
                gen_s=gen_s, w=w)


def cpu_baseline_hulls(w, nsample=16384):
    """The oracle's query::contact + query::distance (the reference algorithms) on the first
    pairs of the same workload, all host threads."""
    from oracle import pyoracle as orc
    from ncollide_amd.math import to_oracle_poses
    rec, pool = w["records"], w["pool"]
    params = np.stack([rec["p0"], rec["p1"], rec["p2"]], 1)
    oss = orc.ShapeSet(rec["type"], params, rec["hull_offset"], rec["hull_count"], pool, dtype=np.float32)
    op = to_oracle_poses(w["poses"])
    ia, ib = w["ia"][:nsample], w["ib"][:nsample]
    threads = min(os.cpu_count() or 1, 64)
    t = time.perf_counter()
    c = orc.contact(oss, op, ia, ib, 0.0, nthreads=threads)
    _, dit = orc.distance(oss, op, ia, ib, nthreads=threads)
    dt = time.perf_counter() - t
    # FP32 work of the support scans: 5 flops (3 mul, 2 add) per hull vertex per support call,
    # support calls = the initial CSO point of each query + one per GJK and EPA iteration
    nv = rec["hull_count"].astype(np.float64)
    calls = 2.0 + c["gjk_iters"] + c["epa_iters"] + dit
    flops_per_pair = float(np.mean(5.0 * (nv[ia] + nv[ib]) * calls))
    n1 = min(2048, len(ia))
    t = time.perf_counter()
    orc.contact(oss, op, ia[:n1], ib[:n1], 0.0, nthreads=1)
    orc.distance(oss, op, ia[:n1], ib[:n1], nthreads=1)
    st = n1 / (time.perf_counter() - t)
    return dict(value=len(ia) / dt, unit="pairs/s (contact + distance)", cores=threads, kind="port",
                sample=f"first {len(ia)} pairs of the same workload, query::contact + query::distance "
                       f"(GJK/EPA) per pair, {threads} threads",
                single_thread_pairs_per_s=st, epa_iters_per_pair=float(c["epa_iters"].mean()),
                flops_per_pair=flops_per_pair)


def bench_learned(args, ncx, torch, ctx):
    """Learned warm start on coherent hull pairs: train the 44-64-64-3 MLP on plain-GJK results
    (the library's own FP16 MFMA kernels), then per step contact + distance on every pair with
    init_dir = None (plain) or the MLP's direction (GJK capped at 32, plain GJK on the cap)."""
    from ncollide_amd import learned
    t = time.perf_counter()
    feat, y = learned.training_set(ctx, seed=41, nsamples=args.train_samples)
    gen_s = time.perf_counter() - t
    mlp, hist, train_s = learned.train(ctx, feat, y, epochs=8)
    # MLP forward throughput (MFMA): one call over the training features
    nf = feat.shape[0]
    fwd = _timed(torch, lambda: mlp.forward(feat), 5)
    flops = 14208.0 * nf
    res = learned.evaluate_coherent(ctx, mlp, seed=4, npairs=args.hull_pairs, steps=args.c4_steps,
                                    variants=("plain", "learned", "temporal", "oracle_dir"))
    p, l, tp, ex = res["plain"], res["learned"], res["temporal"], res["oracle_dir"]
    return dict(pairs=res["pairs"], steps=res["steps"], train_samples=int(nf), train_set_s=gen_s, train_s=train_s,
                _mlp=mlp,
                temporal_pairs_per_s=tp["pairs_per_s"], exact_direction_pairs_per_s=ex["pairs_per_s"],
                temporal_iters={"contact": tp["contact_gjk_iters_per_pair"],
                                "distance": tp["distance_gjk_iters_per_pair"]},
                exact_direction_iters={"contact": ex["contact_gjk_iters_per_pair"],
                                       "distance": ex["distance_gjk_iters_per_pair"]},
                loss_first=hist[0], loss_last=hist[-1],
                plain_pairs_per_s=p["pairs_per_s"], learned_pairs_per_s=l["pairs_per_s"],
                speedup=l["pairs_per_s"] / p["pairs_per_s"],
                plain_iters={"contact": p["contact_gjk_iters_per_pair"], "distance": p["distance_gjk_iters_per_pair"]},
                learned_iters={"contact": l["contact_gjk_iters_per_pair"],
                               "distance": l["distance_gjk_iters_per_pair"]},
                agreement_with_plain={"contact": l["contact_agreement_with_plain"],
                                      "distance": l["distance_agreement_with_plain"]},
                mlp_forward={"samples": int(nf), "ms": fwd * 1e3, "tflops": flops / fwd / 1e12,
                             "frac_of_fp16_dense_peak": flops / fwd / 2.5e15})


def cpu_accuracy_learned(ncx, torch, ctx, mlp, nsample=16384):
    """Error against the f64 oracle (the error reference) of plain and learned-warm-start contact
    depth and distance on the learned workload's hull distribution: relative errors' median and
    p99 (depth on pairs in contact in both, distance on separated pairs)."""
    from oracle import pyoracle as orc
    from ncollide_amd import workloads
    from ncollide_amd.math import to_oracle_poses
    from ncollide_amd.warmstart import descriptors
    w = workloads.hull_pairs(seed=44, npairs=nsample)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=ctx.device)
    dirs = mlp.predict_dirs(ss, poses, descriptors(ss), w["ia"], w["ib"])
    runs = {"plain": (ncx.query.contact(ss, poses, w["ia"], w["ib"], 0.0), ncx.query.distance(ss, poses, w["ia"], w["ib"])[0]),
            "learned": (ncx.query.contact(ss, poses, w["ia"], w["ib"], 0.0, init_dir=dirs, max_iter=32),
                        ncx.query.distance(ss, poses, w["ia"], w["ib"], init_dir=dirs, max_iter=32)[0])}
    rec = w["records"]
    params = np.stack([rec["p0"], rec["p1"], rec["p2"]], 1)
    o64 = orc.ShapeSet(rec["type"], params, rec["hull_offset"], rec["hull_count"], w["pool"], dtype=np.float64)
    op = to_oracle_poses(w["poses"]).astype(np.float64)
    threads = min(os.cpu_count() or 1, 64)
    td, _ = orc.distance(o64, op, w["ia"], w["ib"], nthreads=threads)
    tc = orc.contact(o64, op, w["ia"], w["ib"], 0.0, nthreads=threads)
    out = {"pairs": nsample}
    for key, (c, d) in runs.items():
        d = d.cpu().numpy().astype(np.float64)
        sep = td > 0
        ed = np.abs(d[sep] - td[sep]) / np.maximum(td[sep], 1e-3)
        h = c["has_contact"].cpu().numpy() & tc["has"]
        ec = np.abs(c["depth"].cpu().numpy().astype(np.float64)[h] - tc["depth"][h]) / np.maximum(np.abs(tc["depth"][h]), 1e-3)
        out[key] = {"distance_rel_err_p50": float(np.median(ed)), "distance_rel_err_p99": float(np.percentile(ed, 99)),
                    "depth_rel_err_p50": float(np.median(ec)), "depth_rel_err_p99": float(np.percentile(ec, 99))}
    return out


def bench_scene(args, ncx, torch, ctx, mlp=None):
    """Scene step on 1,048,576 mixed shapes in 64 Gaussian clusters: shape AABBs, LBVH refit
    (rebuilt every --rebuild-every steps, default 2), pair search, contacts; each refit step replayed
    as a hipGraph.
    Then the same steps with the narrow phase off, for the broad phase's own time and roofline:
    bytes per step = 68 N (pose + parameters in, box out) + 72 (N-1) (refit) + 64 x nodes visited
    by one self-query traversal (both child boxes per visit) + 8 x pairs, plus on a rebuild step
    64 N (key/value sort) + 16 (N-1) (hierarchy), spread over the rebuild cycle."""
    from ncollide_amd import workloads
    from ncollide_amd.world import World
    t = time.perf_counter()
    w = workloads.full_pipeline(seed=5, n=args.scene_shapes)
    gen_s = time.perf_counter() - t
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    vel = torch.as_tensor(w["velocities"], device=ctx.device)

    def run(contacts, steps, warm=None):
        poses = torch.as_tensor(w["poses"], device=ctx.device)
        world = World(ss, margin=w["margin"], prediction=0.0, max_pairs=args.scene_max_pairs,
                      rebuild_every=args.rebuild_every)
        world.set_contacts(contacts)
        if warm == "learned":
            world.set_warmstart(mlp)
        elif warm == "temporal":
            world.set_temporal_warmstart(True)
        times, pairs, large = [], [], 0
        for s in range(steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            world.step(poses, use_graph=True)
            npairs, ovf = world.sync()
            times.append(time.perf_counter() - t0)
            pairs.append(npairs)
            large += ovf
            poses[:, 0:3] += vel
        torch.cuda.synchronize()
        return world, times, pairs, large

    world, times, pairs, large = run(True, args.scene_steps)
    ms = float(np.mean(times[1:] if len(times) > 1 else times)) * 1e3
    res = dict(shapes=args.scene_shapes, steps=args.scene_steps, ms_per_step=ms, ms_first_step=times[0] * 1e3,
               first_tier_variant_by_bucket=ss.tuning("contact")[0],
               first_tier_sample_ms_by_bucket=ss.tuning("contact")[1],
               pairs_per_step=float(np.mean(pairs)), pairs_per_s=float(np.sum(pairs[1:])) / float(np.sum(times[1:])),
               epa_pairs_past_first_arena=int(large), gen_s=gen_s)
    del world
    # the same steps with the persistent narrow phase's last_gjk_dir, and with the learned init_dir
    for warm in ("temporal", "learned"):
        if warm == "learned" and mlp is None:
            continue
        world, times, pairs, _ = run(True, args.scene_steps, warm=warm)
        res[f"{warm}_warm_start"] = {"ms_per_step": float(np.mean(times[1:])) * 1e3,
                                     "pairs_per_s": float(np.sum(pairs[1:])) / float(np.sum(times[1:]))}
        del world
    # broad phase alone: steps 1..16 timed
    world, times, pairs, _ = run(False, 17)
    nodes = world.query_stats()
    n = args.scene_shapes
    p = float(np.mean(pairs[1:]))
    per_step = 68 * n + 72 * (n - 1) + 64 * nodes + 8 * p + (64 * n + 16 * (n - 1)) / args.rebuild_every
    bp_s = float(np.mean(times[1:]))
    res["broad_phase"] = {"ms_per_step": bp_s * 1e3, "nodes_visited_per_pass": nodes, "pairs_per_step": p,
                          "roofline": {"bound": "hbm", "achieved": per_step / bp_s / 1e9, "peak": HBM_PEAK / 1e9,
                                       "unit": "GB/s", "frac": per_step / bp_s / HBM_PEAK,
                                       "alg_bytes_per_step": per_step}}
    return res


def bench_plain_pipeline(ncx, torch, ctx, steps=20):
    """One plain-pipeline step: shape AABBs (loosened by the margin), broad-phase proxies + update,
    then query::contact at prediction 0 on every pair."""
    from ncollide_amd import workloads
    w = workloads.plain_pipeline(seed=1)
    dev = ctx.device
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    poses = torch.as_tensor(w["poses"], device=dev)

    def step():
        bb = ss.aabbs(poses, margin=w["margin"])
        bp = ncx.DBVTBroadPhase(w["margin"], ctx=ctx)
        bp.create_proxies(bb.cpu().numpy().T)
        bp.update()
        ia, ib = bp.pairs_tensors()
        g = ncx.query.contact(ss, poses, ia, ib, w["prediction"])
        torch.cuda.synchronize()
        return ia.numel(), int(g["has_contact"].sum().item())

    step()
    t0 = time.perf_counter()
    for _ in range(steps):
        n, c = step()
    dt = (time.perf_counter() - t0) / steps
    return dict(ms_per_step=dt * 1e3, pairs=n, contacts=c)


def cpu_baseline_plain(nsteps=20):
    """The plain pipeline on the reference algorithms (the oracle's DBVTBroadPhase, then
    query::contact on every pair), one thread as the reference runs: ms per step."""
    from oracle import pyoracle as orc
    from ncollide_amd import workloads
    from ncollide_amd.math import to_oracle_poses
    w = workloads.plain_pipeline(seed=1)
    rec = w["records"]
    params = np.stack([rec["p0"], rec["p1"], rec["p2"]], 1)
    oss = orc.ShapeSet(rec["type"], params, rec["hull_offset"], rec["hull_count"], w["pool"], dtype=np.float32)
    op = to_oracle_poses(w["poses"])
    t = time.perf_counter()
    for _ in range(nsteps):
        bb = orc.shape_aabbs(oss, op) + np.array([-1, -1, -1, 1, 1, 1], np.float32) * np.float32(w["margin"])
        bp = orc.BroadPhase(w["margin"])
        for b in bb:
            bp.create_proxy(b)
        bp.update()
        keys = bp.pairs()
        ia = (keys >> np.uint64(32)).astype(np.uint32)
        ib = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        orc.contact(oss, op, ia, ib, 0.0, nthreads=1)
    ms = (time.perf_counter() - t) / nsteps * 1e3
    return dict(value=ms, unit="ms per step", cores=1, kind="port",
                sample="the whole 1,024-shape step (shape AABBs, DBVTBroadPhase update, query::contact per pair)")


def cpu_baseline_scene(ncx, torch, ctx, args, nsample=65536):
    """query::contact on a sample of one scene step's pairs on the oracle, all host threads."""
    from oracle import pyoracle as orc
    from ncollide_amd import workloads
    from ncollide_amd.math import to_oracle_poses
    from ncollide_amd.world import World
    w = workloads.full_pipeline(seed=5, n=args.scene_shapes)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    world = World(ss, margin=w["margin"], prediction=0.0, max_pairs=args.scene_max_pairs)
    world.set_contacts(False)
    world.step(torch.as_tensor(w["poses"], device=ctx.device), use_graph=False)
    world.sync()
    r = world.results()
    ia, ib = r["ia"].cpu().numpy().astype(np.uint32), r["ib"].cpu().numpy().astype(np.uint32)
    rng = np.random.default_rng(0)
    sel = rng.choice(len(ia), min(nsample, len(ia)), replace=False)
    rec = w["records"]
    params = np.stack([rec["p0"], rec["p1"], rec["p2"]], 1)
    oss = orc.ShapeSet(rec["type"], params, rec["hull_offset"], rec["hull_count"], w["pool"], dtype=np.float32)
    threads = min(os.cpu_count() or 1, 64)
    t = time.perf_counter()
    orc.contact(oss, to_oracle_poses(w["poses"]), ia[sel], ib[sel], 0.0, nthreads=threads)
    dt = time.perf_counter() - t
    return dict(value=len(sel) / dt, unit="pairs/s (query::contact)", cores=threads, kind="port",
                sample=f"{len(sel)} random pairs of the first step's {len(ia)}, query::contact per pair")


def bench_queries(args, ncx, torch, ctx):
    """Throughput of the other batched queries on the same workloads: time of impact on the hull
    pairs (with velocities toward each other), ray casts on the scene's shapes, first ray hits
    among the 1,048,576-shape scene's proxies, and contact / distance of shapes against the
    257x257 heightfield TriMesh."""
    from ncollide_amd import workloads, query
    from ncollide_amd.world import World
    out = {}
    dev = ctx.device
    # time of impact: hull pairs of the batched-contact distribution, shape b moving back toward a
    w = workloads.hull_pairs(seed=3, npairs=args.hull_pairs)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    rng = np.random.default_rng(7)
    vel = np.zeros((len(w["records"]), 3), np.float32)
    vel[1::2] = (-w["dirs"] * rng.uniform(0.2, 1.0, (len(w["dirs"]), 1))).astype(np.float32)
    poses = torch.as_tensor(w["poses"], device=dev)
    velt = torch.as_tensor(vel, device=dev)
    ia = torch.as_tensor(w["ia"].astype(np.int32), device=dev)
    ib = torch.as_tensor(w["ib"].astype(np.int32), device=dev)
    t = _timed(torch, lambda: query.time_of_impact(ss, poses, velt, ia, ib, 10.0, 0.0), 5)
    out["time_of_impact"] = {"pairs": len(w["ia"]), "ms": t * 1e3, "pairs_per_s": len(w["ia"]) / t}
    if not args.no_cpu:
        from oracle import pyoracle as orc
        from ncollide_amd.math import to_oracle_poses
        rec = w["records"]
        oss = orc.ShapeSet(rec["type"], np.stack([rec["p0"], rec["p1"], rec["p2"]], 1), rec["hull_offset"],
                           rec["hull_count"], w["pool"], dtype=np.float32)
        k = min(65536, len(w["ia"]))
        nth = min(16, os.cpu_count() or 1)
        tc = time.perf_counter()
        orc.time_of_impact(oss, to_oracle_poses(w["poses"]), vel, w["ia"][:k], w["ib"][:k], np.float32(10.0),
                           np.float32(0.0), nthreads=nth)
        out["time_of_impact"]["cpu_baseline"] = {
            "value": k / (time.perf_counter() - tc), "unit": "pairs/s", "cores": nth, "kind": "port",
            "sample": f"first {k} pairs, query::time_of_impact (GJK ray cast on the Minkowski difference)"}
    # scene shapes: ray casts, world first interference
    sw = workloads.full_pipeline(seed=5, n=args.scene_shapes)
    sss = ncx.ShapeSet(records=sw["records"], pool=sw["pool"], ctx=ctx)
    sp = torch.as_tensor(sw["poses"], device=dev)
    nr = 1 << 20
    idx = rng.integers(0, args.scene_shapes, nr).astype(np.int32)
    c = sw["poses"][idx, :3]
    o = (c + rng.normal(0, 2.0, (nr, 3))).astype(np.float32)
    d = (c + rng.uniform(-0.2, 0.2, (nr, 3)) - o).astype(np.float32)
    ot, dt_, it = (torch.as_tensor(a, device=dev) for a in (o, d, idx))
    t = _timed(torch, lambda: query.cast_rays(sss, sp, it, ot, dt_, 50.0, True), 5)
    out["shape_ray_casts"] = {"rays": nr, "ms": t * 1e3, "rays_per_s": nr / t}
    if not args.no_cpu:
        from oracle import pyoracle as orc
        from ncollide_amd.math import to_oracle_poses
        rec = sw["records"]
        soss = orc.ShapeSet(rec["type"], np.stack([rec["p0"], rec["p1"], rec["p2"]], 1), rec["hull_offset"],
                            rec["hull_count"], sw["pool"], dtype=np.float32)
        k = 262144
        nth = min(16, os.cpu_count() or 1)
        tc = time.perf_counter()
        orc.shape_rays(soss, to_oracle_poses(sw["poses"]), idx[:k].astype(np.uint32), o[:k], d[:k], np.float32(50.0),
                       True, nthreads=nth)
        out["shape_ray_casts"]["cpu_baseline"] = {
            "value": k / (time.perf_counter() - tc), "unit": "rays/s", "cores": nth, "kind": "port",
            "sample": f"first {k} rays, RayCast::toi_and_normal_with_ray of each ray's shape"}
    world = World(sss, margin=sw["margin"], prediction=0.0, max_pairs=args.scene_max_pairs)
    world.set_contacts(False)
    world.step(sp, use_graph=False)
    world.sync()
    ro = np.concatenate([rng.uniform(0, 200, (nr, 2)), np.full((nr, 1), 210.0)], 1).astype(np.float32)
    rd = np.tile(np.array([[0.05, 0.02, -1.0]], np.float32), (nr, 1))
    t = _timed(torch, lambda: world.first_interference_with_rays(sp, ro, rd, 400.0), 3)
    h, _, _ = world.first_interference_with_rays(sp, ro, rd, 400.0)
    out["world_first_interference"] = {"rays": nr, "ms": t * 1e3, "rays_per_s": nr / t,
                                       "hit_fraction": float((h != -1).float().mean().item())}
    del world
    # composite: the heightfield TriMesh against scene shapes hovering over it
    pts, tri = workloads.heightfield_mesh(seed=2)
    mesh = ncx.TriMesh(pts, tri, ctx=ctx)
    nq = 1 << 18
    mp = sw["poses"].copy()  # poses are indexed by shape: the first nq shapes move over the terrain
    mp[:nq, 0:2] = rng.uniform(0.5, 25.1, (nq, 2))
    mp[:nq, 2] = rng.uniform(-0.6, 1.2, nq)
    mpt = torch.as_tensor(mp, device=dev)
    qi = torch.arange(nq, dtype=torch.int32, device=dev)
    tcn = _timed(torch, lambda: mesh.contact_with_shapes(sss, mpt, qi, 0.0), 3)
    g = mesh.contact_with_shapes(sss, mpt, qi, 0.0)
    tdi = _timed(torch, lambda: mesh.distance_to_shapes(sss, mpt, qi), 3)
    out["composite_trimesh"] = {"queries": nq, "triangles": len(tri), "contact_ms": tcn * 1e3,
                                "contact_queries_per_s": nq / tcn, "distance_ms": tdi * 1e3,
                                "distance_queries_per_s": nq / tdi,
                                "contact_fraction": float(g["has_contact"].float().mean().item())}
    if not args.no_cpu:
        from oracle import pyoracle as orc
        from ncollide_amd.math import to_oracle_poses
        rec = sw["records"]
        coss = orc.ShapeSet(rec["type"], np.stack([rec["p0"], rec["p1"], rec["p2"]], 1), rec["hull_offset"],
                            rec["hull_count"], sw["pool"], dtype=np.float32)
        kc, kd = 16384, 4096  # a distance query takes about seven times the CPU time of a contact
        nth = min(16, os.cpu_count() or 1)
        mp7 = np.array([0, 0, 0, 0, 0, 0, 1], np.float32)
        op = to_oracle_poses(mp)
        tc = time.perf_counter()
        orc.mesh_contact(coss, pts, tri, mp7, op, np.arange(kc, dtype=np.uint32), np.float32(0.0), nthreads=nth)
        tcc = time.perf_counter() - tc
        tc = time.perf_counter()
        orc.mesh_distance(coss, pts, tri, mp7, op, np.arange(kd, dtype=np.uint32), nthreads=nth)
        tcd = time.perf_counter() - tc
        out["composite_trimesh"]["cpu_baseline"] = {
            "value": 2.0 / (tcc / kc + tcd / kd), "unit": "queries/s (one contact + one distance per query)",
            "cores": nth, "kind": "port", "contact_queries_per_s": kc / tcc, "distance_queries_per_s": kd / tcd,
            "sample": f"first {kc} queries for contact_composite_shape_shape, first {kd} for "
                      "distance_composite_shape_shape"}
    return out


def bench_manifolds(args, ncx, torch, ctx):
    """Persistent contact manifolds (ncx_contact_manifolds) on the pairs of the 1,048,576-shape
    scene, its hulls drawn so that each has a polyhedral structure (ConvexHull::try_new succeeds):
    one slot per pair, four updates with the shapes moving, prediction 0.02 / 0.05 rad."""
    from ncollide_amd import workloads
    from ncollide_amd.world import World
    dev = ctx.device
    t0 = time.perf_counter()
    sw = workloads.full_pipeline(seed=5, n=args.scene_shapes, polyhedra=True)
    gen_s = time.perf_counter() - t0
    ss = ncx.ShapeSet(records=sw["records"], pool=sw["pool"], ctx=ctx)
    poses = torch.as_tensor(sw["poses"], device=dev)
    world = World(ss, margin=sw["margin"], prediction=0.0, max_pairs=args.scene_max_pairs)
    world.set_contacts(False)
    world.step(poses, use_graph=False)
    world.sync()
    r = world.results()
    ia, ib = r["ia"].clone(), r["ib"].clone()
    del world
    n = ia.numel()
    slots = torch.arange(n, dtype=torch.int32, device=dev)
    gm = ncx.ContactManifolds(ctx, n)
    vel = torch.as_tensor(sw["velocities"] * 0.05, device=dev)
    g = gm.update(ss, poses, ia, ib, slots, 0.02, 0.05, 0.05)  # first update (also builds the hull faces)
    times, counts = [], None
    for step in range(4):
        poses[:, 0:3] += vel
        torch.cuda.synchronize()
        t = time.perf_counter()
        g = gm.update(ss, poses, ia, ib, slots, 0.02, 0.05, 0.05, out=g)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        counts = g["count"]
    tmed = float(np.median(times))
    c = counts.cpu().numpy()
    out = {"pairs": n, "ms_per_update": tmed * 1e3, "pairs_per_s": n / tmed,
           "contacts_per_s": float(c.sum()) / tmed, "contacts_per_manifold": float(c.mean()),
           "manifolds_by_count": np.bincount(np.minimum(c, 9), minlength=10).tolist(), "gen_s": gen_s,
           "data": "1,048,576-shape clustered scene (hulls redrawn where try_new fails), pairs of one broad-phase "
                   "step, four persistent updates"}
    if not args.no_cpu:
        from oracle import pyoracle as orc
        from ncollide_amd.hull import hull_triangles
        from ncollide_amd.math import to_oracle_poses
        rec, pool = sw["records"], sw["pool"]
        k = min(65536, n)
        sel = np.sort(np.random.default_rng(1).choice(n, k, replace=False))
        sa, sb = ia.cpu().numpy()[sel].astype(np.uint32), ib.cpu().numpy()[sel].astype(np.uint32)
        used = np.unique(np.concatenate([sa, sb]))
        tris, toff, tcnt = [], np.zeros(len(rec), np.uint32), np.zeros(len(rec), np.uint32)
        o = 0
        for i in used[rec["type"][used] == ncx.CONVEX_HULL]:
            t = hull_triangles(pool[rec["hull_offset"][i]:rec["hull_offset"][i] + rec["hull_count"][i]])
            tris.append(t)
            toff[i], tcnt[i] = o, len(t)
            o += len(t)
        params = np.stack([rec["p0"], rec["p1"], rec["p2"]], 1)
        oss = orc.ShapeSet(rec["type"], params, rec["hull_offset"], rec["hull_count"], pool, dtype=np.float32)
        store = orc.ManifoldStore(k)
        nth = min(16, os.cpu_count() or 1)
        op = to_oracle_poses(poses.cpu().numpy())
        t = time.perf_counter()
        store.step(oss, np.concatenate(tris), toff, tcnt, op, sa, sb, np.arange(k, dtype=np.uint32), 0.02, 0.05,
                   0.05, nthreads=nth)
        tc = time.perf_counter() - t
        out["cpu_baseline"] = {"value": k / tc, "unit": "pairs/s (one manifold update)", "cores": nth, "kind": "port",
                               "sample": f"{k} random pairs of the step, one update from fresh slots (hull faces "
                                         "built outside the timing)"}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _pmc_pass(exe, counter, args):
    """One rocprofv3 --pmc pass over a short ray-cast run (a child process): per-dispatch values of
    `counter` (KiB) for k_cast_wide and for the 256 MiB calibration copy."""
    d = tempfile.mkdtemp(prefix="ncx_pmc_", dir="/tmp")
    cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable, "-u",
           os.path.abspath(__file__), "--quick", "--no-cpu", "--no-traffic", "--copy-calibration", "--steps", "3",
           "--warmup", "1", "--rays", str(args.rays)]
    p = subprocess.Popen(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), stdout=subprocess.DEVNULL,
                         stderr=subprocess.DEVNULL, start_new_session=True)
    try:
        p.wait(timeout=240)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.wait()
        shutil.rmtree(d, ignore_errors=True)
        return None
    cast, copy = [], []
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith("counter_collection.csv"):
                with open(os.path.join(root, f)) as fh:
                    for r in csv.DictReader(fh):
                        if r.get("Counter_Name") != counter:
                            continue
                        if "k_cast_wide" in r["Kernel_Name"]:
                            cast.append(float(r["Counter_Value"]))
                        elif "copyBuffer" in r["Kernel_Name"]:
                            copy.append(float(r["Counter_Value"]))
    shutil.rmtree(d, ignore_errors=True)
    if not cast or not copy:
        return None
    return {"cast_kib": sum(cast) / len(cast), "dispatches": len(cast), "copy_kib": max(copy)}


def measure_traffic(args):
    """HBM bytes per k_cast_wide launch, read in this run: rocprofv3 FETCH_SIZE, then WRITE_SIZE, in two
    child runs of the cast, each scaled by what the same counter reads for a device-to-device copy of
    256 MiB in that run.  Runs before this process touches the GPU; None when rocprofv3 is absent or a
    pass fails."""
    exe = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3") else None)
    if exe is None:
        return None
    f = _pmc_pass(exe, "FETCH_SIZE", args)
    w = _pmc_pass(exe, "WRITE_SIZE", args) if f else None
    if not f or not w:
        return None
    copy_kib = 256.0 * 1024.0
    ff, fw = copy_kib / f["copy_kib"], copy_kib / w["copy_kib"]
    fetch, write = f["cast_kib"] * ff * 1024.0, w["cast_kib"] * fw * 1024.0
    return {"bytes_per_launch": fetch + write, "fetch_bytes": fetch, "write_bytes": write,
            "fetch_kib_raw": f["cast_kib"], "write_kib_raw": w["cast_kib"], "fetch_factor": ff, "write_factor": fw,
            "dispatches": f["dispatches"], "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes in this run"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is a "
                                   "design note only)", "n_gpus": args.gpus}))
        sys.exit(2)
    traffic = None
    if not args.no_traffic:
        log("ray-cast HBM traffic (rocprofv3 FETCH_SIZE, WRITE_SIZE passes) ...")
        traffic = measure_traffic(args)
    import torch
    import ncollide_amd as ncx
    ctx = ncx.default_context()
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device=ctx.device)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    log("ray casting ...")
    r = bench_rays(args, ncx, torch, ctx)
    value = r["n"] / (r["wall_ms"] / 1e3)
    result = {
        "metric": "convex pairs/sec + rays/sec at 1 GPU; learned-warm-start vs plain GJK @1e-5 tol",
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded heightfield TriMesh and cosine-weighted rays)",
        "config": {"workload": "heightfield_rays: 1,048,576 rays x 257x257 heightfield TriMesh (131,072 triangles), "
                               "max_toi 100, first-hit toi + normal + triangle id",
                   "rays": r["n"], "triangles": 131072, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": r["achieved"] / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": r["achieved"] / HBM_PEAK,
                     "traffic": traffic["bytes_per_launch"] if traffic else _traffic("raycast"),
                     "kernel": "k_cast_wide", "kernel_ms": r["kernel_ms"],
                     "alg_bytes_per_launch": r["alg_bytes"], "nodes_per_ray": r["nodes_per_ray"],
                     "tris_per_ray": r["tris_per_ray"]},
        "hit_fraction": r["hits"],
    }
    if traffic:
        # counter bytes over the kernel time measured in this run
        traffic["hbm_gbs"] = traffic["bytes_per_launch"] / (r["kernel_ms"] / 1e3) / 1e9
        traffic["hbm_frac"] = traffic["hbm_gbs"] * 1e9 / HBM_PEAK
        result["roofline"]["traffic_measured"] = traffic
    else:
        result["roofline"]["traffic_source"] = "profiles/raycast_traffic.json (committed rocprofv3 passes)"
    if not args.no_cpu:
        log("cpu baseline ...")
        cb = cpu_baseline_rays(r["pts"], r["tri"], r["o"], r["d"])
        result["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        result["cpu_baseline"]["single_thread_rays_per_s"] = cb["single_thread"]
        result["cpu_baseline"]["cpu_model"] = cpu_model()
        log("ray accuracy vs f64 ...")
        result["accuracy_vs_f64"] = cpu_accuracy_rays(r)
    if not args.quick:
        sec = {}
        want = (lambda k: args.only is None or args.only == k)
        if want("hulls"):
            log("hull pairs ...")
            h = bench_hulls(args, ncx, torch, ctx)
            sec["hull_pairs"] = {k: v for k, v in h.items() if k not in ("w", "gen_s")}
            if not args.no_cpu:
                cb = cpu_baseline_hulls(h["w"])
                fpp = cb.pop("flops_per_pair")
                sec["hull_pairs"]["cpu_baseline"] = cb
                # the FP32 leg of the roofline: support-scan flops per pair (oracle's iteration
                # counts) over the measured time; roofline pairs/s = 1 / max(bytes/BW, flops/FP32 peak)
                hp = sec["hull_pairs"]
                t_pair = (hp["contact_ms"] + hp["distance_ms"]) / 1e3 / hp["pairs"]
                roof = 1.0 / max(hp["roofline"]["bytes_per_pair"] / HBM_PEAK, fpp / FP32_PEAK)
                hp["roofline"]["valu"] = {"flops_per_pair": fpp, "achieved_tflops": fpp / t_pair / 1e12,
                                          "peak_tflops": FP32_PEAK / 1e12, "frac": fpp / t_pair / FP32_PEAK}
                hp["roofline"]["roofline_pairs_per_s"] = roof
                hp["roofline"]["frac_of_roofline_pairs_per_s"] = (1.0 / t_pair) / roof
        mlp = None
        if want("learned"):
            log("learned warm start ...")
            lw = bench_learned(args, ncx, torch, ctx)
            mlp = lw.pop("_mlp")
            sec["learned_warm_start"] = lw
            if not args.no_cpu:
                log("learned accuracy vs f64 ...")
                lw["accuracy_vs_f64"] = cpu_accuracy_learned(ncx, torch, ctx, mlp)
        if want("scene"):
            log("scene step ...")
            sec["scene_step"] = bench_scene(args, ncx, torch, ctx, mlp)
            if not args.no_cpu:
                sec["scene_step"]["cpu_baseline"] = cpu_baseline_scene(ncx, torch, ctx, args)
        if want("plain"):
            log("plain pipeline ...")
            sec["plain_pipeline"] = bench_plain_pipeline(ncx, torch, ctx)
            if not args.no_cpu:
                sec["plain_pipeline"]["cpu_baseline"] = cpu_baseline_plain()
        if want("queries"):
            log("other queries ...")
            sec["queries"] = bench_queries(args, ncx, torch, ctx)
        if want("manifolds"):
            log("contact manifolds ...")
            try:
                sec["contact_manifolds"] = bench_manifolds(args, ncx, torch, ctx)
            except Exception as e:  # reported, not fatal to the headline line
                sec["contact_manifolds"] = {"error": str(e)}
        result["secondary"] = sec
    print(json.dumps(result))


def _traffic(name):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    factor) for this workload, from the committed profile summary; None when not measured."""
    p = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    return None


if __name__ == "__main__":
    main()
