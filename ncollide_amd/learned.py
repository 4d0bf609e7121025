"""Learned warm start experiment: training data, training, and the coherent-motion evaluation.

Workload (coherent_hulls): the hull-pair distribution of the batched GJK/EPA workload
(262,144 pairs, N~U{8..256} points per hull), moved for 64 steps with per-shape linear
velocity in a ball of radius 0.05/step and angular velocity in a ball of 0.05 rad/step.
Each step runs contact (prediction 0) and distance on every pair.  Variants per step:
  plain     init_dir = None (the reference's default directions)
  learned   init_dir = MLP prediction, GJK capped at 32 iterations, plain GJK on the cap
  temporal  init_dir = the previous step's final GJK/EPA direction (the reference's own warm start,
            last_gjk_dir of ConvexPolyhedronConvexPolyhedronManifoldGenerator,
            src/pipeline/narrow_phase/contact_generator/convex_polyhedron_convex_polyhedron_manifold_generator.rs:96-145)
  oracle-dir init_dir = this step's final GJK/EPA direction (what a perfect predictor reaches)
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


def _contact_dirs(ss, poses, ia, ib):
    """Per-pair final GJK/EPA direction (the contact normal at a large prediction, so separated
    pairs get their closest-points direction too); zero rows where there is none."""
    g = query.contact(ss, poses, ia, ib, 10.0)
    return torch.where(g["has_contact"][:, None], g["normal"], torch.zeros_like(g["normal"])).contiguous()


def evaluate_coherent(ctx, mlp, seed=4, npairs=1 << 18, steps=64, placer_iters=20, log=None, cap=32,
                      variants=("plain", "learned", "temporal", "oracle_dir")):
    """Per step and variant: contact (prediction 0) and distance on every pair, both from the
    variant's init_dir (GJK capped at `cap`, plain GJK on the cap).  Reports GJK iterations per
    pair for each query, pairs/s for the two queries together, and agreement with plain."""
    placer, st = _placer(ctx)
    w = workloads.hull_pairs(seed=seed, npairs=npairs, placer=placer, bisect_iters=placer_iters)
    ss, ia, ib = st["ss"], st["ia"], st["ib"]
    dev = ctx.device
    rng = np.random.default_rng(seed + 1000)
    v, om = workloads.coherent_motion(rng, 2 * npairs)
    poses_np = w["poses"]
    desc = descriptors(ss)
    keys = tuple(variants)
    assert keys[0] == "plain"
    res = {k: dict(citers=0.0, diters=0.0, time=0.0, agree=0.0, dagree=0.0) for k in keys}
    prev_dirs = None
    n = len(w["ia"])

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    def run(poses, dirs_fn):
        def go():
            dirs = dirs_fn()
            if dirs is None:
                c = query.contact(ss, poses, ia, ib, 0.0)
                d = query.distance(ss, poses, ia, ib)
            else:
                c = query.contact(ss, poses, ia, ib, 0.0, init_dir=dirs, max_iter=cap)
                d = query.distance(ss, poses, ia, ib, init_dir=dirs, max_iter=cap)
            return c, d
        return timed(go)

    for s in range(steps):
        poses = torch.as_tensor(poses_np, device=dev)
        need_dirs = "temporal" in keys or "oracle_dir" in keys
        this_dirs = _contact_dirs(ss, poses, ia, ib) if need_dirs else None
        out = {}
        out["plain"] = run(poses, lambda: None)
        if "learned" in keys:
            out["learned"] = run(poses, lambda: mlp.predict_dirs(ss, poses, desc, ia, ib))
        if "temporal" in keys:
            out["temporal"] = out["plain"] if prev_dirs is None else run(poses, lambda: prev_dirs)
        if "oracle_dir" in keys:
            out["oracle_dir"] = run(poses, lambda: this_dirs)
        (pc, (pd, _)), _ = out["plain"]
        for key in keys:
            (c, (d, dit)), t = out[key]
            r = res[key]
            r["citers"] += c["gjk_iters"].float().mean().item() / steps
            r["diters"] += dit.float().mean().item() / steps
            r["time"] += t
            r["agree"] += _agree(c, pc) / steps
            r["dagree"] += ((d - pd).abs() <= 1e-5 * pd.abs() + 1e-6).float().mean().item() / steps
        prev_dirs = this_dirs
        poses_np = workloads.step_poses(poses_np, v, om)
        if log and s % 16 == 0 and "learned" in keys:
            (c0, (d0, i0)), _ = out["plain"]
            (c1, (d1, i1)), _ = out["learned"]
            log(f"step {s}: contact GJK plain {c0['gjk_iters'].float().mean().item():.2f} / learned "
                f"{c1['gjk_iters'].float().mean().item():.2f} it; distance GJK plain {i0.float().mean().item():.2f}"
                f" / learned {i1.float().mean().item():.2f} it")
    for key in keys:
        r = res[key]
        r["pairs_per_s"] = n * steps / r["time"]
        r["contact_gjk_iters_per_pair"] = r.pop("citers")
        r["distance_gjk_iters_per_pair"] = r.pop("diters")
        r["contact_agreement_with_plain"] = r.pop("agree")
        r["distance_agreement_with_plain"] = r.pop("dagree")
        r["seconds"] = r.pop("time")
    if "learned" in keys:
        res["speedup_learned_vs_plain"] = res["learned"]["pairs_per_s"] / res["plain"]["pairs_per_s"]
    res["pairs"] = n
    res["steps"] = steps
    return res
