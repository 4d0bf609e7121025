"""Learned warm start experiment: training data, training, and the coherent-motion evaluation.

Workload (coherent_hulls): the hull-pair distribution of the batched GJK/EPA workload
(262,144 pairs, N~U{8..256} points per hull), moved for 64 steps with per-shape linear
velocity in a ball of radius 0.05/step and angular velocity in a ball of 0.05 rad/step.
Variants per step:
  plain     init_dir = None (the reference's default direction, t2 - t1)
  learned   init_dir = MLP prediction, GJK capped at 32 iterations, plain GJK on the cap
  temporal  init_dir = the previous step's contact normal (the reference's own warm start,
            last_gjk_dir of ConvexPolyhedronConvexPolyhedronManifoldGenerator,
            src/pipeline/narrow_phase/contact_generator/convex_polyhedron_convex_polyhedron_manifold_generator.rs:96-145)
  oracle-dir init_dir = this step's plain result normal (the bound a perfect predictor reaches)
"""
import time

import numpy as np
import torch

from . import query, workloads
from .shape import SHAPE_DTYPE, CONVEX_HULL, ShapeSet
from .warmstart import WarmStartMLP, descriptors, features, frame1_targets


def _placer(ctx):
    state = {}

    def placer(records, pool, ia, ib):
        dev = ctx.device
        state["ss"] = ShapeSet(records=records, pool=pool, ctx=ctx)
        state["ia"] = torch.as_tensor(ia.astype(np.int32), device=dev)
        state["ib"] = torch.as_tensor(ib.astype(np.int32), device=dev)

        def sd(poses8):
            g = query.contact(state["ss"], torch.as_tensor(poses8, device=dev), state["ia"], state["ib"], 10.0)
            s = (-g["depth"]).cpu().numpy()
            s[~g["has_contact"].cpu().numpy()] = 10.0
            return s
        return sd
    return placer, state


def training_set(ctx, seed=41, nsamples=2_000_000, pool_hulls=65536, bisect_iters=20):
    """(features fp16 (n,48), targets (n,3) in shape-1 frame) from plain GJK/EPA on random pairs of
    hulls drawn from a pool, placed like the hull-pair workload (half penetrating, half separated)."""
    rng = np.random.default_rng(seed)
    pts, offs, cnts = workloads.ellipsoid_clouds(rng, pool_hulls)
    pool, hoff, hcnt = workloads.build_hulls(pts, offs, cnts)
    n = nsamples
    pick = rng.integers(0, pool_hulls, 2 * n)
    rec = np.zeros(2 * n, SHAPE_DTYPE)
    rec["type"] = CONVEX_HULL
    rec["hull_offset"] = hoff[pick]
    rec["hull_count"] = hcnt[pick]
    q = workloads.uniform_quats(rng, 2 * n)
    u = workloads.unit_vectors(rng, n)
    target = np.where(np.arange(n) % 2 == 0, -rng.uniform(0, 0.1, n), rng.uniform(0, 0.5, n))
    ia = np.arange(0, 2 * n, 2, dtype=np.uint32)
    ib = ia + 1
    placer, st = _placer(ctx)
    sd = placer(rec, pool, ia, ib)
    # bisection of the translation along u (same rule as workloads.hull_pairs)
    ext = np.full(n, 2.5)
    lo, hi = np.zeros(n), ext
    t = np.zeros((2 * n, 3), np.float32)
    for _ in range(bisect_iters):
        mid = 0.5 * (lo + hi)
        t[1::2] = (u * mid[:, None]).astype(np.float32)
        s = sd(workloads.poses8(t, q))
        above = s > target
        hi = np.where(above, mid, hi)
        lo = np.where(above, lo, mid)
    t[1::2] = (u * (0.5 * (lo + hi))[:, None]).astype(np.float32)
    poses = torch.as_tensor(workloads.poses8(t, q), device=ctx.device)
    ss = st["ss"]
    g = query.contact(ss, poses, st["ia"], st["ib"], 10.0)
    keep = g["has_contact"]
    desc = descriptors(ss)
    feat = features(ss, poses, desc, st["ia"], st["ib"])
    y = frame1_targets(poses, st["ia"], g["normal"])
    return feat[keep].contiguous(), y[keep].contiguous()


def train(ctx, feat, y, epochs=8, batch=8192, lr=2e-3, seed=0, log=None):
    mlp = WarmStartMLP(ctx, seed=seed)
    t = time.perf_counter()
    hist = mlp.fit(feat, y, epochs=epochs, batch=batch, lr=lr, log=log)
    torch.cuda.synchronize()
    return mlp, hist, time.perf_counter() - t


def _agree(a, b):
    """Fraction of pairs on which two contact results agree to 1e-5 rel + 1e-6 abs in depth and
    1e-4 rad in normal (and on has_contact)."""
    ha, hb = a["has_contact"], b["has_contact"]
    same = ha == hb
    both = ha & hb
    da, db = a["depth"][both], b["depth"][both]
    okd = (da - db).abs() <= 1e-5 * db.abs() + 1e-6
    na, nb = a["normal"][both], b["normal"][both]
    cos = (na * nb).sum(1) / (na.norm(dim=1) * nb.norm(dim=1)).clamp_min(1e-30)
    okn = torch.arccos(cos.clamp(-1, 1)) <= 1e-4
    bad = (~same).sum().item() + (~(okd & okn)).sum().item()
    return 1.0 - bad / ha.numel()


def evaluate_coherent(ctx, mlp, seed=4, npairs=1 << 18, steps=64, placer_iters=20, log=None):
    placer, st = _placer(ctx)
    w = workloads.hull_pairs(seed=seed, npairs=npairs, placer=placer, bisect_iters=placer_iters)
    ss, ia, ib = st["ss"], st["ia"], st["ib"]
    dev = ctx.device
    rng = np.random.default_rng(seed + 1000)
    v, om = workloads.coherent_motion(rng, 2 * npairs)
    poses_np = w["poses"]
    desc = descriptors(ss)
    res = {k: dict(iters=0.0, time=0.0, agree=0.0) for k in ("plain", "learned", "temporal", "oracle_dir")}
    prev_normal = None
    n = len(w["ia"])

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    for s in range(steps):
        poses = torch.as_tensor(poses_np, device=dev)
        plain, tp = timed(lambda: query.contact(ss, poses, ia, ib, 0.0))

        def learned_fn():
            dirs = mlp.predict_dirs(ss, poses, desc, ia, ib)
            return query.contact(ss, poses, ia, ib, 0.0, init_dir=dirs, max_iter=32)
        learned, tl = timed(learned_fn)
        if prev_normal is None:
            temporal, tt = plain, tp
        else:
            temporal, tt = timed(lambda: query.contact(ss, poses, ia, ib, 0.0, init_dir=prev_normal))
        od = torch.where(plain["has_contact"][:, None], plain["normal"],
                         torch.zeros_like(plain["normal"])).contiguous()
        odir, to = timed(lambda: query.contact(ss, poses, ia, ib, 0.0, init_dir=od))
        for key, r, t in (("plain", plain, tp), ("learned", learned, tl), ("temporal", temporal, tt),
                          ("oracle_dir", odir, to)):
            res[key]["iters"] += r["gjk_iters"].float().mean().item() / steps
            res[key]["time"] += t
            res[key]["agree"] += _agree(r, plain) / steps
        prev_normal = torch.where(plain["has_contact"][:, None], plain["normal"],
                                  torch.zeros_like(plain["normal"])).contiguous()
        poses_np = workloads.step_poses(poses_np, v, om)
        if log and s % 16 == 0:
            log(f"step {s}: plain {plain['gjk_iters'].float().mean().item():.2f} it, "
                f"learned {learned['gjk_iters'].float().mean().item():.2f} it")
    for key in res:
        res[key]["pairs_per_s"] = n * steps / res[key]["time"]
        res[key]["gjk_iters_per_pair"] = res[key].pop("iters")
        res[key]["agreement_with_plain"] = res[key].pop("agree")
        res[key]["seconds"] = res[key].pop("time")
    res["pairs"] = n
    res["steps"] = steps
    return res
