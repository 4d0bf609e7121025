"""CPU tests that pin the oracle (the C++ restatement in oracle/) to the reference's
own known answers and to what mathematics fixes: closed forms, brute force and
invariants.  A plausible restatement mistake (a dropped term, a wrong sign or
index, a swapped operand) fails at least one of these."""
import itertools

import numpy as np
import pytest

from ncollide_amd import workloads


def ident(n=1):
    return np.tile(np.array([0, 0, 0, 0, 0, 0, 1.0]), (n, 1))


def pose7(t, q=(0, 0, 0, 1)):
    return np.array(list(t) + list(q), float)


def shapeset(orc, shapes, dtype=np.float64):
    """shapes: list of ('ball', r) / ('cuboid', (hx,hy,hz)) / ('capsule', hh, r) / ('hull', pts)."""
    types, params, off, cnt, pool = [], [], [], [], []
    o = 0
    for s in shapes:
        if s[0] == "ball":
            types.append(0); params.append([s[1], 0, 0]); off.append(0); cnt.append(0)
        elif s[0] == "cuboid":
            types.append(1); params.append(list(s[1])); off.append(0); cnt.append(0)
        elif s[0] == "capsule":
            types.append(2); params.append([s[1], s[2], 0]); off.append(0); cnt.append(0)
        else:
            p = np.asarray(s[1], float).reshape(-1, 3)
            types.append(3); params.append([0, 0, 0]); off.append(o); cnt.append(len(p)); pool.append(p)
            o += len(p)
    pool = np.concatenate(pool) if pool else np.zeros((0, 3))
    return orc.ShapeSet(types, params, off, cnt, pool, dtype=dtype)


# ------------------------------------------------------------------ reference known answers
def test_epa3_cuboid_known_answer(orc):
    """build/ncollide3d/tests/geometry/epa3.rs:7-22 (f64, exact equality)."""
    ss = shapeset(orc, [("cuboid", (2, 1, 1)), ("cuboid", (2, 1, 1))])
    r = orc.contact_support_map(ss, np.stack([pose7((3.5, 0, 0)), pose7((0, 0, 0))]), [0], [1], 10.0)
    assert r["has"][0]
    assert r["depth"][0] == 0.5
    assert np.array_equal(r["normal"][0], [-1.0, 0.0, 0.0])
    r = orc.contact_support_map(ss, np.stack([pose7((0, 0.2, 0)), pose7((0, 0, 0))]), [0], [1], 10.0)
    assert r["depth"][0] == 1.8
    assert np.array_equal(r["normal"][0], [0.0, -1.0, 0.0])


@pytest.mark.parametrize("k", [1, 2, 3])
def test_contact_query3d_example(orc, k):
    """build/ncollide3d/examples/contact_query3d.rs: ball r=1 vs cube h=1 at (k,k,k), prediction 1.
    Closed form: depth = 1 - sqrt(3) (k - 1)."""
    ss = shapeset(orc, [("ball", 1.0), ("cuboid", (1, 1, 1))])
    r = orc.contact(ss, np.stack([pose7((k, k, k)), pose7((0, 0, 0))]), [0], [1], 1.0)
    expect = 1 - np.sqrt(3) * (k - 1)
    if k == 3:
        assert not r["has"][0]
    else:
        assert r["has"][0]
        assert abs(r["depth"][0] - expect) < 1e-12
        assert (r["depth"][0] > 0) == (k == 1)


def test_dbvt_broad_phase_example(orc):
    """build/ncollide3d/examples/dbvt_broad_phase3d.rs:21-61: 6 interferences, then 1 after removing two."""
    bp = orc.BroadPhase(0.2)
    hs = [bp.create_proxy([x - .5, y - .5, -.5, x + .5, y + .5, .5]) for x, y in [(0, 0), (0, .5), (.5, 0), (.5, .5)]]
    bp.update()
    assert bp.num_interferences() == 6
    bp.remove(hs[:2])
    bp.update()
    assert bp.num_interferences() == 1


def test_touching_cuboids_no_nan(orc):
    """build/ncollide3d/tests/geometry/contact.rs:8-26 (#182): cuboids (0.5,0.5,0.1) touching at y = 1."""
    ss = shapeset(orc, [("cuboid", (0.5, 0.5, 0.1)), ("cuboid", (0.5, 0.5, 0.1))])
    r = orc.contact(ss, np.stack([pose7((0, 0, 0)), pose7((0, 1.0, 0))]), [0], [1], 0.0)
    assert np.all(np.isfinite(r["depth"])) and np.all(np.isfinite(r["normal"]))


def test_coincident_cuboids_axis_normal(orc):
    """build/ncollide3d/tests/pipeline/contact_pairs.rs:9-85: coincident unit cuboids -> GJK from the
    +x fallback ends on a segment through the origin and EPA starts from a triangle through it.
    A contact exists, its normal is axis-parallel, and its depth is 0 (EPA keeps the face with the
    smallest support distance, epa3.rs:344-355, here the initial triangle) -- the reference test
    skips exactly this `contact.depth == 0.0` case."""
    ss = shapeset(orc, [("cuboid", (1, 1, 1)), ("cuboid", (1, 1, 1))])
    r = orc.contact(ss, np.stack([pose7((0, 0, 0)), pose7((0, 0, 0))]), [0], [1], 0.0)
    assert r["has"][0] and r["depth"][0] == 0.0
    assert np.sum(np.abs(r["normal"][0]) == 1.0) == 1
    # Offset by 0.5 along y: EPA takes its numerical-issue early return (epa3.rs:397-402) on this
    # degenerate start; the contact still exists with a finite, unit normal.
    r = orc.contact(ss, np.stack([pose7((0, 0.5, 0)), pose7((0, 0, 0))]), [0], [1], 0.0)
    assert r["has"][0] and 0 < r["depth"][0] <= 1.5 and abs(np.linalg.norm(r["normal"][0]) - 1) < 1e-12


def test_triangle_aabb_under_rotation(orc):
    """src/bounding_volume/aabb_triangle.rs:44-71 pins the triangle AABB to the support-map AABB.
    The hull AABB (point_cloud_aabb) of a triangle's 3 points at a pi/2 rotation must equal the
    support-map AABB of the same points (6 support calls)."""
    pts = np.array([[-1, 2, 3], [4, -5, 6], [7, 8, -9.0]])
    q = (np.sin(np.pi / 4), 0, 0, np.cos(np.pi / 4))
    ss = shapeset(orc, [("hull", pts)])
    bb = orc.shape_aabbs(ss, pose7((1, 2, 3), q)[None])[0]
    w = np.array([_rot(q, p) + [1, 2, 3] for p in pts])
    assert np.allclose(bb[:3], w.min(0), atol=1e-12) and np.allclose(bb[3:], w.max(0), atol=1e-12)


def _rot(q, v):
    qv, w = np.array(q[:3]), q[3]
    t = 2 * np.cross(qv, v)
    return np.asarray(v) + w * t + np.cross(qv, t)


# ------------------------------------------------------------------ Voronoi projections vs brute force
def _closest_on_simplex_brute(P):
    """Exact closest point of the convex hull of P to the origin (minimisation over every face
    of the simplex, each solved by least squares with barycentric clamping)."""
    best = None
    n = len(P)
    for k in range(1, n + 1):
        for sub in itertools.combinations(range(n), k):
            S = P[list(sub)]
            if k == 1:
                c = S[0]
            else:
                A = (S[1:] - S[0]).T
                lam, *_ = np.linalg.lstsq(A, -S[0], rcond=None)
                if np.any(lam < -1e-12) or lam.sum() > 1 + 1e-12:
                    continue
                c = S[0] + A @ lam
            d = np.linalg.norm(c)
            if best is None or d < best[0]:
                best = (d, c)
    return best


@pytest.mark.parametrize("kind", [2, 3, 4])
def test_voronoi_projection_matches_brute_force(orc, kind):
    rng = np.random.default_rng(10 + kind)
    for _ in range(300):
        P = rng.normal(size=(kind, 3)) + rng.normal(size=3) * 0.7
        proj, lk, li, bc = orc.project_origin(P)
        if lk == 3:  # origin inside the tetrahedron
            assert kind == 4
            assert np.allclose(proj, 0)
            lam = np.linalg.solve(np.vstack([P.T, np.ones(4)]), np.array([0, 0, 0, 1.0]))
            assert np.all(lam >= -1e-9)
            continue
        d, c = _closest_on_simplex_brute(P)
        assert np.linalg.norm(proj - c) < 1e-9 * max(1, np.abs(P).max())


def test_benches_simplex_fixture(orc):
    """build/ncollide3d/benches/query/algorithm.rs:5-23: the tetrahedron fixture of the simplex bench."""
    P = np.array([[-.5, -.5, -.5], [0, .5, 0], [.5, -.5, -.5], [0, -.5, -.5]])
    proj, lk, li, bc = orc.project_origin(P)
    d, c = _closest_on_simplex_brute(P)
    assert np.linalg.norm(proj - c) < 1e-12


# ------------------------------------------------------------------ ray casting
def test_ray_triangle_closed_form(orc):
    rng = np.random.default_rng(3)
    for _ in range(200):
        tri = rng.normal(size=(3, 3))
        o = rng.normal(size=3) * 3
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        ok, toi, n, fid = orc.ray_triangle(tri, o, d)
        # Moller-Trumbore in f64 as the independent closed form.
        e1, e2 = tri[1] - tri[0], tri[2] - tri[0]
        p = np.cross(d, e2)
        det = e1 @ p
        s = o - tri[0]
        u = s @ p / det
        q = np.cross(s, e1)
        v = d @ q / det
        t = e2 @ q / det
        hit = u >= 0 and v >= 0 and u + v <= 1 and t >= 0
        if abs(u) < 1e-9 or abs(v) < 1e-9 or abs(u + v - 1) < 1e-9:
            continue
        assert ok == hit
        if hit:
            assert abs(toi - t) < 1e-9 * max(1, t)
            nn = np.cross(e1, e2)
            nn /= np.linalg.norm(nn)
            assert abs(abs(n @ nn) - 1) < 1e-12 and n @ d <= 0  # faces the ray origin
            assert fid == (0 if nn @ d < 0 else 1)


def test_ray_aabb_slab_closed_form(orc):
    bb = np.array([-1, -2, -3, 1, 2, 3.0])
    ok, t = orc.ray_aabb(bb, [-5, 0, 0], [1, 0, 0], 100.0)
    assert ok and t == 4.0
    ok, t = orc.ray_aabb(bb, [-5, 5, 0], [1, 0, 0], 100.0)
    assert not ok
    ok, t = orc.ray_aabb(bb, [0, 0, 0], [0, 0, 1], 100.0, solid=True)
    assert ok and t == 0.0
    ok, t = orc.ray_aabb(bb, [0, 0, 0], [0, 0, 1], 100.0, solid=False)
    assert ok and t == 3.0


def test_trimesh_best_first_equals_brute_force(orc):
    pts, tri = workloads.heightfield_mesh(seed=2, cells=32)
    o, d = workloads.heightfield_rays(seed=2, n=3000, extent=3.2, height=1.0)
    m = orc.TriMesh(pts, tri)
    toi, n, f, visits = m.cast(o, d, 100.0)
    bt, bn, bf, nh = m.brute(o, d, 100.0)
    ties = 0
    for k in range(len(o)):
        if f[k] != bf[k]:
            ties += 1
            assert toi[k] == bt[k]  # only exact toi ties may pick a different triangle
    assert np.array_equal(toi, bt)
    assert ties <= 2
    hits = np.mean(f != 0xFFFFFFFF)
    assert 0.2 < hits < 1.0  # rays leaving the small patch sideways miss it


# ------------------------------------------------------------------ GJK / EPA against exact answers
def _sat_depth(A, B, ta, tb):
    """Exact penetration depth of two convex polytopes (vertex sets A, B placed at ta, tb, identity
    rotation): minimum over candidate axes (face normals of both and edge x edge) of the overlap."""
    from scipy.spatial import ConvexHull
    Aw, Bw = A + ta, B + tb
    axes = []
    for P in (Aw, Bw):
        h = ConvexHull(P)
        axes += list(h.equations[:, :3])
    ea = _edges(Aw)
    eb = _edges(Bw)
    for u in ea:
        for v in eb:
            c = np.cross(u, v)
            nc = np.linalg.norm(c)
            if nc > 1e-9:
                axes.append(c / nc)
    best = np.inf
    for ax in axes:
        pa, pb = Aw @ ax, Bw @ ax
        ov = min(pa.max() - pb.min(), pb.max() - pa.min())
        best = min(best, ov)
    return best


def _edges(P):
    from scipy.spatial import ConvexHull
    h = ConvexHull(P)
    es = set()
    for s in h.simplices:
        for i in range(3):
            a, b = sorted((s[i], s[(i + 1) % 3]))
            es.add((a, b))
    return [P[b] - P[a] for a, b in es]


def _exact_distance(Aw, Bw):
    """Exact distance between convex polytopes: min over vertex-face, face-vertex and edge-edge pairs."""
    from scipy.spatial import ConvexHull
    best = np.inf
    for P, Q in ((Aw, Bw), (Bw, Aw)):
        h = ConvexHull(Q)
        for s in h.simplices:
            T = Q[s]
            for p in P:
                _, c = _closest_on_simplex_brute(T - p)
                best = min(best, np.linalg.norm(c))
    for (a0, a1) in _edge_pts(Aw):
        for (b0, b1) in _edge_pts(Bw):
            best = min(best, _seg_seg(a0, a1, b0, b1))
    return best


def _edge_pts(P):
    from scipy.spatial import ConvexHull
    h = ConvexHull(P)
    es = set()
    for s in h.simplices:
        for i in range(3):
            es.add(tuple(sorted((s[i], s[(i + 1) % 3]))))
    return [(P[a], P[b]) for a, b in es]


def _seg_seg(p0, p1, q0, q1):
    best = np.inf
    for s in np.linspace(0, 1, 1):
        pass
    d1, d2, r = p1 - p0, q1 - q0, p0 - q0
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c, b = d1 @ r, d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0, 1) if den > 1e-15 else 0.0
    t = (b * s + f) / e
    if t < 0:
        t, s = 0.0, np.clip(-c / a, 0, 1)
    elif t > 1:
        t, s = 1.0, np.clip((b - c) / a, 0, 1)
    return np.linalg.norm(p0 + d1 * s - (q0 + d2 * t))


def test_gjk_distance_exact_for_polytopes(orc):
    rng = np.random.default_rng(7)
    for _ in range(25):
        A = rng.normal(size=(10, 3)) * rng.uniform(0.3, 1.0, 3)
        B = rng.normal(size=(10, 3)) * rng.uniform(0.3, 1.0, 3)
        u = workloads.unit_vectors(rng, 1)[0]
        ta, tb = np.zeros(3), u * rng.uniform(3.0, 5.0)
        ss = shapeset(orc, [("hull", A), ("hull", B)])
        d, it = orc.distance(ss, np.stack([pose7(ta), pose7(tb)]), [0], [1])
        ex = _exact_distance(A + ta, B + tb)
        if ex > 0:
            assert abs(d[0] - ex) < 1e-6 * max(1, ex)


def test_warm_started_distance_rules(orc):
    """init_dir = the exact separating direction keeps the exact distance and needs fewer GJK
    iterations on average (not on every pair: the first support point along the exact normal is
    a vertex of the closest face, and GJK still has to assemble that face); a cap of one
    iteration from a poor direction falls back to plain GJK, so the distance equals the plain one
    bit for bit and the count is 1 + the plain count."""
    rng = np.random.default_rng(17)
    plain_its, warm_its = [], []
    for _ in range(25):
        A = rng.normal(size=(40, 3)) * rng.uniform(0.3, 1.0, 3)
        B = rng.normal(size=(40, 3)) * rng.uniform(0.3, 1.0, 3)
        u = workloads.unit_vectors(rng, 1)[0]
        tb = u * rng.uniform(3.0, 5.0)
        ss = shapeset(orc, [("hull", A), ("hull", B)])
        poses = np.stack([pose7((0, 0, 0)), pose7(tb)])
        d0, it0 = orc.distance(ss, poses, [0], [1])
        if d0[0] <= 0:  # penetrating: _exact_distance measures boundaries only
            continue
        ex = _exact_distance(A, B + tb)
        # exact direction from shape 1 to shape 2 (the contact normal at a large prediction)
        r = orc.contact(ss, poses, [0], [1], 100.0)
        n = r["normal"][0]
        d1, it1 = orc.distance(ss, poses, [0], [1], init_dirs=n[None], max_iter=32)
        assert abs(d1[0] - ex) < 1e-6 * max(1, ex)
        plain_its.append(it0[0])
        warm_its.append(it1[0])
        d2, it2 = orc.distance(ss, poses, [0], [1], init_dirs=-n[None], max_iter=1)
        assert d2[0] == d0[0]
        assert it2[0] == 1 + it0[0]
        # an all-zero direction is None
        d3, it3 = orc.distance(ss, poses, [0], [1], init_dirs=np.zeros((1, 3)), max_iter=1)
        assert d3[0] == d0[0] and it3[0] == it0[0]
    assert len(plain_its) > 10 and np.mean(warm_its) < np.mean(plain_its)


def test_epa_depth_exact_for_polytopes(orc):
    rng = np.random.default_rng(8)
    checked = 0
    for _ in range(40):
        A = rng.normal(size=(12, 3)) * 0.7
        B = rng.normal(size=(12, 3)) * 0.7
        tb = rng.normal(size=3) * 0.3
        depth_sat = _sat_depth(A, B, np.zeros(3), tb)
        if depth_sat <= 1e-6:
            continue
        ss = shapeset(orc, [("hull", A), ("hull", B)])
        r = orc.contact(ss, np.stack([pose7((0, 0, 0)), pose7(tb)]), [0], [1], 0.0)
        assert r["has"][0]
        assert abs(r["depth"][0] - depth_sat) < 1e-6 * max(1, depth_sat)
        # invariant: depth = -n . (w2 - w1)
        assert abs(r["depth"][0] + r["normal"][0] @ (r["world2"][0] - r["world1"][0])) < 1e-12
        checked += 1
    assert checked > 20


def test_ball_ball_and_ball_cuboid_closed_forms(orc):
    rng = np.random.default_rng(9)
    for _ in range(100):
        r1, r2 = rng.uniform(0.1, 1, 2)
        c2 = rng.normal(size=3) * 1.5
        ss = shapeset(orc, [("ball", r1), ("ball", r2)])
        r = orc.contact(ss, np.stack([pose7((0, 0, 0)), pose7(c2)]), [0], [1], 0.5)
        dist = np.linalg.norm(c2)
        assert r["has"][0] == (dist < r1 + r2 + 0.5)
        if r["has"][0]:
            assert abs(r["depth"][0] - (r1 + r2 - dist)) < 1e-12
        d, _ = orc.distance(ss, np.stack([pose7((0, 0, 0)), pose7(c2)]), [0], [1])
        assert abs(d[0] - max(0.0, dist - r1 - r2)) < 1e-12
    for _ in range(100):
        r1 = rng.uniform(0.1, 1)
        h = rng.uniform(0.2, 1, 3)
        c1 = rng.normal(size=3) * 1.5
        ss = shapeset(orc, [("ball", r1), ("cuboid", h)])
        r = orc.contact(ss, np.stack([pose7(c1), pose7((0, 0, 0))]), [0], [1], 10.0)
        outside = np.maximum(np.abs(c1) - h, 0)
        if np.any(outside > 0):
            expect = r1 - np.linalg.norm(outside)
        else:
            expect = r1 + np.min(h - np.abs(c1))
        assert abs(r["depth"][0] - expect) < 1e-12


def test_capsule_support_distance(orc):
    """Capsule vs ball distance: closed form |c - closest point on the segment| - r_c - r_b."""
    rng = np.random.default_rng(11)
    for _ in range(50):
        hh, rc, rb = rng.uniform(0.1, 0.5), rng.uniform(0.05, 0.3), rng.uniform(0.1, 0.5)
        c = rng.normal(size=3) * 2 + np.array([0, 0, 2.0])
        ss = shapeset(orc, [("capsule", hh, rc), ("ball", rb)])
        d, _ = orc.distance(ss, np.stack([pose7((0, 0, 0)), pose7(c)]), [0], [1])
        y = np.clip(c[1], -hh, hh)
        ex = np.linalg.norm(c - np.array([0, y, 0])) - rc - rb
        if ex > 1e-3:
            assert abs(d[0] - ex) < 2e-3 * max(1, ex)  # GJK's own relative stopping rule on curved shapes


def test_broad_phase_matches_brute_force(orc):
    w = workloads.plain_pipeline(seed=1)
    ss = orc.ShapeSet(w["records"]["type"], np.stack([w["records"]["p0"], w["records"]["p1"], w["records"]["p2"]], 1),
                      np.zeros(1024), np.zeros(1024), np.zeros((0, 3)), dtype=np.float32)
    from ncollide_amd.math import to_oracle_poses
    bb = orc.shape_aabbs(ss, to_oracle_poses(w["poses"]))
    bb[:, :3] -= 0.02
    bb[:, 3:] += 0.02
    bp = orc.BroadPhase(0.02)
    for b in bb:
        bp.create_proxy(b)
    bp.update()
    assert np.array_equal(bp.pairs(), orc.pairs_brute(bb))
    assert 200 < bp.num_interferences() < 450


# ------------------------------------------------------------------ shape AABBs against closed forms
def _quat_mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_shape_aabbs_closed_forms(orc):
    """aabb_ball.rs:8-13 (centre +- r), aabb_cuboid.rs:7-14 (the box of the 8 world corners, which
    centre +- |R| h equals) and aabb_support_map.rs:35-45 on a capsule (the box of the two world
    segment end points, grown by the radius), in f64 with random rotations.  A cuboid box with R in
    place of |R|, or a capsule box missing the radius or the rotation, fails."""
    rng = np.random.default_rng(21)
    q = workloads.uniform_quats(rng, 60).astype(np.float64)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    t = rng.normal(size=(60, 3)) * 5
    for i in range(60):
        r, h, hh = rng.uniform(0.1, 1.0), rng.uniform(0.1, 1.0, 3), rng.uniform(0.1, 1.0)
        ss = shapeset(orc, [("ball", r), ("cuboid", h), ("capsule", hh, r)])
        poses = np.stack([pose7(t[i], q[i])] * 3)
        bb = orc.shape_aabbs(ss, poses)
        assert np.array_equal(bb[0], np.concatenate([t[i] - r, t[i] + r]))
        M = _quat_mat(q[i])
        corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * h
        w = corners @ M.T + t[i]
        assert np.allclose(bb[1], np.concatenate([w.min(0), w.max(0)]), rtol=0, atol=1e-12)
        ends = np.array([[0, hh, 0], [0, -hh, 0]]) @ M.T + t[i]
        assert np.allclose(bb[2], np.concatenate([ends.min(0) - r, ends.max(0) + r]), rtol=0, atol=1e-12)


# ------------------------------------------------------------------ DBVTBroadPhase stored-box rule
def _f32_boxes(rng, n, spread=6.0):
    c = rng.uniform(0, spread, (n, 3)).astype(np.float32)
    e = rng.uniform(0.1, 0.6, (n, 3)).astype(np.float32)
    return np.concatenate([c - e, c + e], 1).astype(np.float32)


def _brute_keys(stored):
    hs = sorted(stored)
    keys = []
    for i, a in enumerate(hs):
        for b in hs[i + 1:]:
            A, B = stored[a], stored[b]
            if np.all(A[:3] <= B[3:]) and np.all(B[:3] <= A[3:]):
                keys.append((min(a, b) << 32) | max(a, b))
    return set(keys)


def test_stored_box_rule_independent_recurrence(orc):
    """dbvt_broad_phase.rs:275-280 (create_proxy stores the box as given) and :325-347
    (deferred_set_bounding_volume stores box.loosened(margin) only when the stored box does not
    contain the new one, aabb.rs:161-163, 190-199), restated here in numpy f32 and run beside the
    oracle over moving steps, a removal and re-created proxies (slab handles reused LIFO).  Stored
    boxes must agree bit for bit (a proxy created since the last update has no stored box, so a
    deferred_set_bounding_volume on it always stores the loosened box); pairs must equal brute
    force over the stored boxes; started /
    stopped events must be the reference's: pairs of removed proxies stop inside remove, the rest
    by set difference at update."""
    rng = np.random.default_rng(31)
    margin = np.float32(0.05)
    bp = orc.BroadPhase(float(margin))
    stored = {}
    for b in _f32_boxes(rng, 120):
        h = bp.create_proxy(b)
        stored[h] = b.copy()
    bp.update()
    prev = _brute_keys(stored)
    st, sp = bp.events()
    assert set(st.tolist()) == prev and len(sp) == 0 and len(prev) > 50
    for step in range(8):
        stopped_at_remove = set()
        detached = set()
        if step in (3, 6):
            gone = sorted(rng.choice(sorted(stored), 15, replace=False).tolist())
            bp.remove(gone)
            for h in gone:
                del stored[h]
            stopped_at_remove = {k for k in prev if (k >> 32) in gone or (k & 0xFFFFFFFF) in gone}
            for b in _f32_boxes(rng, 10):  # re-created proxies take the freed handles, last freed first
                h = bp.create_proxy(b)
                assert h in gone
                stored[h] = b.copy()
                detached.add(h)
        for h in sorted(stored):
            if rng.random() < 0.5:
                cur = stored[h]
                c = (cur[:3] + cur[3:]) * np.float32(0.5) + rng.normal(0, 0.04, 3).astype(np.float32)
                e = rng.uniform(0.1, 0.6, 3).astype(np.float32) if rng.random() < 0.2 else (cur[3:] - cur[:3]) * 0.45
                b = np.concatenate([c - e, c + e]).astype(np.float32)
                bp.deferred_set_bounding_volume(h, b)
                # a proxy created since the last update has no stored box yet: always requeued
                if h in detached or not (np.all(cur[:3] <= b[:3]) and np.all(b[3:] <= cur[3:])):
                    stored[h] = np.concatenate([b[:3] + (-margin), b[3:] + margin]).astype(np.float32)
        bp.update()
        for h, b in stored.items():
            assert np.array_equal(bp.stored_bv(h), b), f"step {step}: stored box of {h}"
        now = _brute_keys(stored)
        assert set(bp.pairs().tolist()) == now
        st, sp = bp.events()
        kept = prev - stopped_at_remove
        assert sorted(st.tolist()) == sorted(now - kept), f"step {step}: started"
        assert sorted(sp.tolist()) == sorted(stopped_at_remove | (kept - now)), f"step {step}: stopped"
        prev = now


# ------------------------------------------------------------------ EPA's 1-simplex start and the triangle test
def test_orthonormal_subspace_first_is_orthogonal(orc):
    """The direction EPA adds its third point along from a segment (nalgebra's
    orthonormal_subspace_basis, epa3.rs:295-305) is orthogonal to the segment and as long as it
    (a unit vector crossed with it), on both branches (|x| > |y| and not)."""
    rng = np.random.default_rng(41)
    for _ in range(400):
        v = rng.normal(size=3) * 10 ** rng.uniform(-3, 3)
        r = orc.orthonormal_subspace_first(v)
        nv = np.linalg.norm(v)
        assert abs(r @ v) <= 1e-13 * nv * nv
        assert abs(np.linalg.norm(r) - nv) <= 1e-13 * nv


def test_is_affinely_dependent_triangle_threshold(orc):
    """utils/triangle.rs:85-102: relative_eq!(|p1p2 x p1p3|^2, 0, epsilon = (100 eps)^2) with
    approx's default max_relative (eps) holds exactly when |p1p2 x p1p3|^2 <= (100 eps)^2 (the
    relative test |x| <= eps |x| never holds for x > 0).  Checked in f64 and f32 on triangles
    scaled across the threshold (cases within 0.1% of it are skipped) and on collinear points."""
    rng = np.random.default_rng(47)
    for dt, eps in ((np.float64, np.finfo(np.float64).eps), (np.float32, np.finfo(np.float32).eps)):
        thr = (100 * float(eps)) ** 2
        tested = [0, 0]
        for _ in range(600):
            p = rng.normal(size=(3, 3))
            c2 = np.sum(np.cross(p[1] - p[0], p[2] - p[0]) ** 2)
            s = (thr / c2) ** 0.25 * 10 ** rng.uniform(-0.25, 0.25)  # |cross|^2 grows as s^4
            p = (p * s).astype(dt)
            x = np.sum(np.cross((p[1] - p[0]).astype(np.float64), (p[2] - p[0]).astype(np.float64)) ** 2)
            if abs(x / thr - 1) < 1e-3:
                continue
            got = orc.is_affinely_dependent_triangle(p[0], p[1], p[2], dtype=dt)
            assert got == (x <= thr), (dt, x, thr)
            tested[int(got)] += 1
        assert min(tested) > 100
        p0, d = rng.normal(size=3), rng.normal(size=3)
        assert orc.is_affinely_dependent_triangle(p0, p0 + d, p0 + 2 * d, dtype=dt)
