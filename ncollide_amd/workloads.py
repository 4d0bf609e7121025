"""Seeded synthetic inputs for the five workloads of the hot path.

  plain_pipeline   1,024 shapes (Ball / Cuboid), broad phase margin 0.02, contact at prediction 0
  heightfield_rays 257x257 heightfield TriMesh (131,072 triangles) x 1,048,576 rays, max_toi 100
  hull_pairs       262,144 ConvexHull pairs, N~U{8..256} points on jittered ellipsoids
  coherent_hulls   the hull pairs moved over 64 coherent steps (learned warm start)
  full_pipeline    1,048,576 mixed shapes in 64 Gaussian clusters, 32 steps

Random streams are numpy PCG64 generators seeded per workload.  Rotations are
uniform unit quaternions (Shoemake), stored (i, j, k, w) like nalgebra.
"""
import numpy as np

from .shape import SHAPE_DTYPE, BALL, CUBOID, CAPSULE, CONVEX_HULL


def uniform_quats(rng, n):
    """Shoemake's uniform random unit quaternions, (n, 4) as (i, j, k, w) float32."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], 1)
    return q.astype(np.float32)


def poses8(t, q):
    out = np.zeros((len(t), 8), np.float32)
    out[:, 0:3] = t
    out[:, 4:8] = q
    return out


def unit_vectors(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


# ------------------------------------------------------------------ plain pipeline (1,024 balls + cuboids)
def plain_pipeline(seed=1, n=1024):
    rng = np.random.default_rng(seed)
    rec = np.zeros(n, SHAPE_DTYPE)
    half = n // 2
    rec["type"][:half] = BALL
    rec["p0"][:half] = rng.uniform(0.1, 0.5, half)
    rec["type"][half:] = CUBOID
    he = rng.uniform(0.1, 0.5, (n - half, 3))
    rec["p0"][half:], rec["p1"][half:], rec["p2"][half:] = he[:, 0], he[:, 1], he[:, 2]
    t = rng.uniform(0.0, 20.0, (n, 3)).astype(np.float32)
    q = uniform_quats(rng, n)
    return dict(records=rec, pool=np.zeros((0, 3), np.float32), poses=poses8(t, q), margin=0.02, prediction=0.0)


# ------------------------------------------------------------------ heightfield TriMesh + rays
def heightfield_mesh(seed=2, cells=256, spacing=0.1):
    """(points (V,3), indices (T,3)): z = sum of 4 sinusoids + U[-0.01, 0.01] noise.
    Sinusoid k runs along x for even k and along y for odd k.  Cell (row r, col c)
    holds triangles 2(r*cells + c) = (p00, p10, p01) and +1 = (p10, p11, p01), CCW seen from +z."""
    rng = np.random.default_rng(seed)
    nv = cells + 1
    amp = rng.uniform(0.05, 0.5, 4)
    freq = rng.uniform(0.5, 4.0, 4)
    phase = rng.uniform(0.0, 2 * np.pi, 4)
    xs = np.arange(nv) * spacing
    X, Y = np.meshgrid(xs, xs)  # Y[r, c] = r * spacing, X[r, c] = c * spacing
    Z = np.zeros_like(X)
    for k in range(4):
        u = X if k % 2 == 0 else Y
        Z += amp[k] * np.sin(freq[k] * u + phase[k])
    Z += rng.uniform(-0.01, 0.01, Z.shape)
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    r, c = np.meshgrid(np.arange(cells), np.arange(cells), indexing="ij")
    p00 = (r * nv + c).ravel()
    p10 = p00 + 1
    p01 = p00 + nv
    p11 = p01 + 1
    tri = np.empty((cells * cells * 2, 3), np.uint32)
    tri[0::2] = np.stack([p00, p10, p01], 1)
    tri[1::2] = np.stack([p10, p11, p01], 1)
    return pts, tri


def heightfield_rays(seed=2, n=1 << 20, extent=25.6, height=5.0):
    """Origins uniform over the footprint at z = height; directions cosine-weighted on the
    downward hemisphere (Malley: r = sqrt(u1), phi = 2 pi u2)."""
    rng = np.random.default_rng(seed + 1000)
    o = np.empty((n, 3), np.float32)
    o[:, 0] = rng.uniform(0, extent, n)
    o[:, 1] = rng.uniform(0, extent, n)
    o[:, 2] = height
    u1, u2 = rng.random(n), rng.random(n)
    r = np.sqrt(u1)
    phi = 2 * np.pi * u2
    d = np.stack([r * np.cos(phi), r * np.sin(phi), -np.sqrt(1.0 - u1)], 1).astype(np.float32)
    return o, d


# ------------------------------------------------------------------ hulls
def ellipsoid_clouds(rng, nclouds, nmin=8, nmax=256, axes=(0.2, 1.0), jitter=0.02):
    counts = rng.integers(nmin, nmax + 1, nclouds).astype(np.uint32)
    offsets = np.zeros(nclouds, np.uint32)
    offsets[1:] = np.cumsum(counts)[:-1]
    total = int(counts.sum())
    v = unit_vectors(rng, total)
    ax = rng.uniform(axes[0], axes[1], (nclouds, 3))
    ax_per_pt = np.repeat(ax, counts, axis=0)
    jit = 1.0 + rng.uniform(-jitter, jitter, (total, 1))
    pts = (v * ax_per_pt * jit).astype(np.float32)
    return pts, offsets, counts


def build_hulls(pts, offsets, counts):
    """Hull vertex pools of many clouds: (pool (P,3) float32, hull_offset, hull_count)."""
    from .hull import convex_hull_batch
    ids, hc = convex_hull_batch(pts, offsets, counts)
    if np.any(hc < 4):
        raise RuntimeError("degenerate hull in the synthetic input")
    sel = _gather_ids(ids, offsets, hc)
    pool = pts[sel]
    hoff = np.zeros(len(hc), np.uint32)
    hoff[1:] = np.cumsum(hc)[:-1]
    return np.ascontiguousarray(pool), hoff, hc


def _gather_ids(ids, offsets, hc):
    total = int(hc.sum())
    starts = np.repeat(offsets.astype(np.int64), hc)
    local = np.arange(total) - np.repeat(np.cumsum(hc) - hc, hc)
    return starts + ids[(starts + local).astype(np.int64)].astype(np.int64)


def hull_pairs(seed=3, npairs=1 << 18, nmin=8, nmax=256, placer=None, bisect_iters=24):
    """Pairs (2k, 2k+1) of ConvexHulls with random rotations.  Pair k is moved along a random
    direction u to a target signed distance: -U[0, 0.1] (penetrating) for even k, U[0, 0.5]
    (separated) for odd k.  `placer(records, pool, ia, ib)` returns a function
    signed_distance(poses) -> (npairs,) evaluating every pair's current signed distance;
    placement bisects the translation along u against it.  Without a placer, pairs are placed
    where their support planes along u touch, shifted by the target."""
    rng = np.random.default_rng(seed)
    nh = 2 * npairs
    pts, offs, cnts = ellipsoid_clouds(rng, nh, nmin, nmax)
    pool, hoff, hcnt = build_hulls(pts, offs, cnts)
    rec = np.zeros(nh, SHAPE_DTYPE)
    rec["type"] = CONVEX_HULL
    rec["hull_offset"] = hoff
    rec["hull_count"] = hcnt
    q = uniform_quats(rng, nh)
    u = unit_vectors(rng, npairs)
    ia = np.arange(0, nh, 2, dtype=np.uint32)
    ib = ia + 1
    signed_distance = placer(rec, pool, ia, ib) if placer is not None else None
    target = np.where(np.arange(npairs) % 2 == 0, -rng.uniform(0.0, 0.1, npairs), rng.uniform(0.0, 0.5, npairs))
    # Separation along u at which the support planes touch (upper bound of the touching distance).
    ext = _support_extent(pool, hoff, hcnt, q, u)
    t0 = ext + target
    t = np.zeros((nh, 3), np.float32)
    if signed_distance is None:
        t[1::2] = (u * t0[:, None]).astype(np.float32)
    else:
        lo = np.zeros(npairs)
        hi = np.maximum(ext + 1.0, 1e-3)
        for _ in range(bisect_iters):
            mid = 0.5 * (lo + hi)
            t[1::2] = (u * mid[:, None]).astype(np.float32)
            sd = np.asarray(signed_distance(poses8(t, q)), np.float64)
            above = sd > target
            hi = np.where(above, mid, hi)
            lo = np.where(above, lo, mid)
        t[1::2] = (u * (0.5 * (lo + hi))[:, None]).astype(np.float32)
    return dict(records=rec, pool=pool, poses=poses8(t, q), ia=ia, ib=ib, target=target, dirs=u)


def _rotate(q, v):
    qv = q[:, :3]
    w = q[:, 3:4]
    t = 2.0 * np.cross(qv, v)
    return v + w * t + np.cross(qv, t)


def _support_extent(pool, hoff, hcnt, q, u):
    """h_A(u) + h_B(-u) per pair for hulls 2k, 2k+1 at the origin with rotations q."""
    npairs = len(u)
    qa, qb = q[0::2].astype(np.float64), q[1::2].astype(np.float64)
    conj = np.array([-1, -1, -1, 1.0])
    la = _rotate(qa * conj, u)
    lb = _rotate(qb * conj, -u)
    out = np.zeros(npairs)
    owner = np.repeat(np.arange(2 * npairs), hcnt)
    d = np.empty(len(pool))
    even = owner % 2 == 0
    pair = owner // 2
    d[even] = np.einsum("ij,ij->i", pool[even].astype(np.float64), la[pair[even]])
    d[~even] = np.einsum("ij,ij->i", pool[~even].astype(np.float64), lb[pair[~even]])
    ha = np.full(2 * npairs, -np.inf)
    np.maximum.at(ha, owner, d)
    out = ha[0::2] + ha[1::2]
    return out


# ------------------------------------------------------------------ coherent motion (learned warm start)
def quat_mul(a, b):
    ai, aj, ak, aw = a[:, 0], a[:, 1], a[:, 2], a[:, 3]
    bi, bj, bk, bw = b[:, 0], b[:, 1], b[:, 2], b[:, 3]
    return np.stack([aw * bi + ai * bw + aj * bk - ak * bj,
                     aw * bj - ai * bk + aj * bw + ak * bi,
                     aw * bk + ai * bj - aj * bi + ak * bw,
                     aw * bw - ai * bi - aj * bj - ak * bk], 1)


def ball_samples(rng, n, radius):
    v = unit_vectors(rng, n)
    r = radius * rng.random(n) ** (1.0 / 3.0)
    return v * r[:, None]


def coherent_motion(rng, n, lin=0.05, ang=0.05):
    return ball_samples(rng, n, lin), ball_samples(rng, n, ang)


def step_poses(poses, v, w):
    """t += v; q <- exp(w / 2) (x) q, renormalised."""
    out = poses.copy()
    out[:, 0:3] += v.astype(np.float32)
    ang = np.linalg.norm(w, axis=1)
    half = 0.5 * ang
    s = np.where(ang > 0, np.sin(half) / np.maximum(ang, 1e-30), 0.5)
    dq = np.concatenate([w * s[:, None], np.cos(half)[:, None]], 1)
    q = quat_mul(dq, poses[:, 4:8].astype(np.float64))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    out[:, 4:8] = q.astype(np.float32)
    return out


# ------------------------------------------------------------------ full pipeline (mixed shapes in clusters)
def polyhedral_hulls(rng, nclouds, nmin=8, nmax=64, max_rounds=20):
    """Hull pools like build_hulls(ellipsoid_clouds(...)) where every hull has a polyhedral
    structure: a cloud whose ConvexHull::try_new fails (merged nearly coplanar faces break
    Euler's characteristic; the reference's try_from_points returns None for it) is drawn
    again.  Returns (pool, hull_offset, hull_count)."""
    from .hull import polyhedra_ok
    pts, offs, cnts = ellipsoid_clouds(rng, nclouds, nmin, nmax)
    clouds = [pts[o:o + c] for o, c in zip(offs, cnts)]
    for _ in range(max_rounds):
        cnt = np.array([len(c) for c in clouds], np.uint32)
        off = np.zeros(len(cnt), np.uint32)
        off[1:] = np.cumsum(cnt)[:-1]
        pool, hoff, hcnt = build_hulls(np.concatenate(clouds), off, cnt)
        ok = polyhedra_ok(pool, hoff, hcnt)
        if ok.all():
            return pool, hoff, hcnt
        bad = np.nonzero(~ok)[0]
        p2, o2, c2 = ellipsoid_clouds(rng, len(bad), nmin, nmax)
        for j, b in enumerate(bad):
            clouds[b] = p2[o2[j]:o2[j] + c2[j]]
    raise RuntimeError("hull clouds kept failing ConvexHull::try_new")


def full_pipeline(seed=5, n=1 << 20, nclusters=64, sigma=5.0, box=200.0, hull_nmax=64, polyhedra=False):
    """Mixed shapes in Gaussian clusters.  polyhedra: every hull has a polyhedral structure (the
    contact-manifold generators need one; see polyhedral_hulls)."""
    rng = np.random.default_rng(seed)
    types = rng.integers(0, 4, n).astype(np.uint32)
    rec = np.zeros(n, SHAPE_DTYPE)
    rec["type"] = types
    nb = types == BALL
    rec["p0"][nb] = rng.uniform(0.1, 0.5, nb.sum())
    nc = types == CUBOID
    he = rng.uniform(0.1, 0.5, (nc.sum(), 3))
    rec["p0"][nc], rec["p1"][nc], rec["p2"][nc] = he[:, 0], he[:, 1], he[:, 2]
    ncap = types == CAPSULE
    rec["p0"][ncap] = rng.uniform(0.1, 0.5, ncap.sum())
    rec["p1"][ncap] = rng.uniform(0.05, 0.3, ncap.sum())
    nh = types == CONVEX_HULL
    hidx = np.nonzero(nh)[0]
    if polyhedra:
        pool, hoff, hcnt = polyhedral_hulls(rng, len(hidx), 8, hull_nmax)
    else:
        pts, offs, cnts = ellipsoid_clouds(rng, len(hidx), 8, hull_nmax)
        pool, hoff, hcnt = build_hulls(pts, offs, cnts)
    rec["hull_offset"][hidx] = hoff
    rec["hull_count"][hidx] = hcnt
    centres = rng.uniform(0, box, (nclusters, 3))
    which = rng.integers(0, nclusters, n)
    t = np.clip(centres[which] + rng.normal(0, sigma, (n, 3)), 0, box).astype(np.float32)
    q = uniform_quats(rng, n)
    vel = ball_samples(rng, n, 0.5).astype(np.float32)
    return dict(records=rec, pool=pool, poses=poses8(t, q), velocities=vel, margin=0.02, prediction=0.0)


# ------------------------------------------------------------------ resting stacks (contact manifolds)
def resting_stacks(seed=7, ncols=16, height=4, spacing=3.0, hull_share=0.25):
    """Columns of shapes resting on a ground cuboid, each a little into the one below, with
    small tilts: face-on-face contacts whose manifolds have several points.  Shape 0 is the
    ground (a cuboid of half extents (ncols * spacing / 2 + 2, ncols * spacing / 2 + 2, 0.5),
    top face at z = 0).  Columns stand on a ncols x ncols grid; each level is a cuboid (a hull
    for hull_share of them) or, at the top, a ball or a lying capsule."""
    rng = np.random.default_rng(seed)
    ncol = ncols * ncols
    n = 1 + ncol * height
    rec = np.zeros(n, SHAPE_DTYPE)
    half = ncols * spacing / 2 + 2.0
    rec["type"][0] = CUBOID
    rec["p0"][0], rec["p1"][0], rec["p2"][0] = half, half, 0.5
    t = np.zeros((n, 3), np.float32)
    q = np.zeros((n, 4), np.float32)
    q[:, 3] = 1.0
    t[0] = (half - 2.0 - spacing / 2, half - 2.0 - spacing / 2, -0.5)
    gx, gy = np.meshgrid(np.arange(ncols), np.arange(ncols), indexing="ij")
    cx, cy = (gx.ravel() * spacing).astype(np.float64), (gy.ravel() * spacing).astype(np.float64)
    types = np.full((ncol, height), CUBOID, np.uint32)
    hullmask = rng.random((ncol, height)) < hull_share
    types[hullmask] = CONVEX_HULL
    top = rng.integers(0, 3, ncol)  # the top level: 0 cuboid/hull as drawn, 1 ball, 2 capsule
    types[top == 1, height - 1] = BALL
    types[top == 2, height - 1] = CAPSULE
    nh = int((types == CONVEX_HULL).sum())
    # hulls: boxes with slightly perturbed corners plus a few points on the faces (flat faces)
    hpts, hoff, hcnt = [], np.zeros(nh, np.uint32), np.zeros(nh, np.uint32)
    he_h = rng.uniform(0.4, 0.9, (nh, 3))
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float64)
    from .hull import convex_hull_vertices, polyhedra_ok
    off = 0
    for i in range(nh):
        for _ in range(50):
            c = corners * he_h[i] * (1 + rng.uniform(-0.01, 0.01, (8, 3)))
            extra = rng.uniform(-1, 1, (6, 3)) * he_h[i] * 0.7
            cl = np.concatenate([c, extra]).astype(np.float32)
            ids = convex_hull_vertices(cl)
            hp = cl[ids]
            if polyhedra_ok(hp, np.zeros(1, np.uint32), np.array([len(hp)], np.uint32))[0]:
                break
        hpts.append(hp)
        hoff[i], hcnt[i] = off, len(hp)
        off += len(hp)
    pool = np.concatenate(hpts).astype(np.float32) if nh else np.zeros((0, 3), np.float32)
    hi = 0
    idx = 1
    for c in range(ncol):
        z = 0.0
        for lv in range(height):
            ty = types[c, lv]
            rec["type"][idx] = ty
            tilt = rng.normal(0, 0.02, 3)
            tilt[2] = rng.uniform(-np.pi, np.pi)  # yaw about z
            if ty == CUBOID:
                he = rng.uniform(0.4, 0.9, 3)
                rec["p0"][idx], rec["p1"][idx], rec["p2"][idx] = he
                hz = he[2]
            elif ty == CONVEX_HULL:
                rec["hull_offset"][idx], rec["hull_count"][idx] = hoff[hi], hcnt[hi]
                hz = float(np.max(np.abs(pool[hoff[hi]:hoff[hi] + hcnt[hi], 2])))
                hi += 1
            elif ty == BALL:
                r = rng.uniform(0.3, 0.7)
                rec["p0"][idx] = r
                hz = r
            else:  # a capsule lying along x: rotate its local y axis onto x
                hh, r = rng.uniform(0.3, 0.8), rng.uniform(0.2, 0.5)
                rec["p0"][idx], rec["p1"][idx] = hh, r
                hz = r
                tilt = np.array([0.0, 0.0, -np.pi / 2]) + rng.normal(0, 0.02, 3)
            pen = rng.uniform(0.0, 0.04)
            zc = z + hz - pen
            t[idx] = (cx[c] + rng.normal(0, 0.05), cy[c] + rng.normal(0, 0.05), zc)
            # rotation: small tilts about x and y, then the yaw (axis-angle composition)
            ang = np.linalg.norm(tilt)
            ax = tilt / ang if ang > 0 else np.array([0.0, 0.0, 1.0])
            q[idx, :3] = ax * np.sin(ang / 2)
            q[idx, 3] = np.cos(ang / 2)
            z = zc + hz
            idx += 1
    vel = rng.normal(0, 0.01, (n, 3)).astype(np.float32)
    vel[0] = 0.0
    return dict(records=rec, pool=pool, poses=poses8(t, q.astype(np.float32)), velocities=vel)
