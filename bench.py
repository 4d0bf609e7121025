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
import json
import os
import sys
import time

import ctypes as C

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # bytes/s, MI355X HBM3E


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--quick", action="store_true", help="skip the secondary workloads")
    p.add_argument("--only", choices=["hulls", "learned", "scene", "plain"], default=None,
                   help="run one secondary workload alone (after the ray-cast headline); for profiling")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--copy-calibration", action="store_true",
                   help="copy 256 MiB device to device once before the timed casts (the reference dispatch "
                        "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
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
    return dict(n=n, wall_ms=wall_ms, kernel_ms=float(np.mean(kms)), kernel_ms_min=float(np.min(kms)),
                alg_bytes=alg_bytes, nodes_per_ray=nodes / n, tris_per_ray=tris / n,
                achieved=alg_bytes / avg_kernel_s, hits=hits, pts=pts, tri=tri, o=o, d=d)


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
                lanes_per_pair_by_bucket={"contact": ss.tuning("contact")[0], "distance": ss.tuning("distance")[0]},
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
    orc.distance(oss, op, ia, ib, nthreads=threads)
    dt = time.perf_counter() - t
    n1 = min(2048, len(ia))
    t = time.perf_counter()
    orc.contact(oss, op, ia[:n1], ib[:n1], 0.0, nthreads=1)
    orc.distance(oss, op, ia[:n1], ib[:n1], nthreads=1)
    st = n1 / (time.perf_counter() - t)
    return dict(value=len(ia) / dt, unit="pairs/s (contact + distance)", cores=threads, kind="port",
                sample=f"first {len(ia)} pairs of the same workload, query::contact + query::distance "
                       f"(GJK/EPA) per pair, {threads} threads",
                single_thread_pairs_per_s=st, epa_iters_per_pair=float(c["epa_iters"].mean()))


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
                                    variants=("plain", "learned"))
    p, l = res["plain"], res["learned"]
    return dict(pairs=res["pairs"], steps=res["steps"], train_samples=int(nf), train_set_s=gen_s, train_s=train_s,
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


def bench_scene(args, ncx, torch, ctx):
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

    def run(contacts, steps):
        poses = torch.as_tensor(w["poses"], device=ctx.device)
        world = World(ss, margin=w["margin"], prediction=0.0, max_pairs=args.scene_max_pairs,
                      rebuild_every=args.rebuild_every)
        world.set_contacts(contacts)
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
               lanes_per_pair_by_bucket=ss.tuning("contact")[0],
               pairs_per_step=float(np.mean(pairs)), pairs_per_s=float(np.sum(pairs[1:])) / float(np.sum(times[1:])),
               epa_pairs_past_first_arena=int(large), gen_s=gen_s)
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


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is a "
                                   "design note only)", "n_gpus": args.gpus}))
        sys.exit(2)
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
                     "frac": r["achieved"] / HBM_PEAK, "traffic": _traffic("raycast"),
                     "kernel": "k_cast_pt", "kernel_ms": r["kernel_ms"],
                     "alg_bytes_per_launch": r["alg_bytes"], "nodes_per_ray": r["nodes_per_ray"],
                     "tris_per_ray": r["tris_per_ray"]},
        "hit_fraction": r["hits"],
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        cb = cpu_baseline_rays(r["pts"], r["tri"], r["o"], r["d"])
        result["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        result["cpu_baseline"]["single_thread_rays_per_s"] = cb["single_thread"]
    if not args.quick:
        sec = {}
        want = (lambda k: args.only is None or args.only == k)
        if want("hulls"):
            log("hull pairs ...")
            h = bench_hulls(args, ncx, torch, ctx)
            sec["hull_pairs"] = {k: v for k, v in h.items() if k not in ("w", "gen_s")}
            if not args.no_cpu:
                sec["hull_pairs"]["cpu_baseline"] = cpu_baseline_hulls(h["w"])
        if want("learned"):
            log("learned warm start ...")
            sec["learned_warm_start"] = bench_learned(args, ncx, torch, ctx)
        if want("scene"):
            log("scene step ...")
            sec["scene_step"] = bench_scene(args, ncx, torch, ctx)
        if want("plain"):
            log("plain pipeline ...")
            sec["plain_pipeline"] = bench_plain_pipeline(ncx, torch, ctx)
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
