"""CPU tests that pin the oracle's contact-manifold generators (the narrow phase's persistent
contacts: polygonal clipping, the capsule-as-segment dispatch, ContactManifold's tracking) and the
hull polyhedral structure they read, to what geometry fixes:

  * a box resting on the ground: its four bottom corners, at the penetration depth;
  * a square on a square turned 45 degrees: the eight corners of their octagon;
  * a ball on a box: one contact, ball feature Face(0) and the box's top face;
  * a capsule lying on the ground: its two end points, depth r - gap, the end caps' features;
    standing: one point; the prediction grows by the radius (a gap below the radius touches);
  * two aligned boxes: coincident clip points collapse to four by the 0.02 tracking distance;
  * ConvexHull::try_new: the library's host restatement equals the oracle's, Euler's
    characteristic holds, faces are planar and every point lies behind every face.

A planted mistake (no clipping, no preprocessor, no capsule prediction, no tracking merge) fails
one of these."""
import numpy as np
import pytest

from ncollide_amd import hull as nh
from tests.test_oracle import shapeset

TOL = 2e-5


def quat_axis(axis, ang):
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    return np.concatenate([a * np.sin(ang / 2), [np.cos(ang / 2)]])


def pose(t, q=(0, 0, 0, 1)):
    return np.array(list(t) + list(q), np.float32)


def hull_tris(shapes):
    tris, off, cnt = [], [], []
    o = 0
    for s in shapes:
        if s[0] == "hull":
            t = nh.hull_triangles(np.asarray(s[1], np.float32))
            tris.append(t)
            off.append(o)
            cnt.append(len(t))
            o += len(t)
        else:
            off.append(0)
            cnt.append(0)
    tr = np.concatenate(tris) if tris else np.zeros((1, 3), np.uint32)
    return tr, np.array(off, np.uint32), np.array(cnt, np.uint32)


def manifold(orc, shapes, poses, pairs, linear=0.0, ang1=0.0, ang2=0.0, store=None, slots=None):
    ss = shapeset(orc, shapes, dtype=np.float32)
    tr, off, cnt = hull_tris(shapes)
    pairs = np.asarray(pairs, np.uint32).reshape(-1, 2)
    if store is None:
        store = orc.ManifoldStore(len(pairs))
    if slots is None:
        slots = np.arange(len(pairs), dtype=np.uint32)
    r = store.step(ss, tr, off, cnt, np.stack(poses), pairs[:, 0], pairs[:, 1], slots, linear, ang1, ang2)
    return r, store


def contacts(r, k):
    n = int(r["count"][k])
    return dict(w1=r["world1"][k, :n], w2=r["world2"][k, :n], n=r["normal"][k, :n], depth=r["depth"][k, :n],
                f1=r["feature1"][k, :n], f2=r["feature2"][k, :n])


GROUND = ("cuboid", (5.0, 5.0, 0.5))
GROUND_POSE = pose((0, 0, -0.5))


@pytest.mark.parametrize("yaw", [0.0, 0.5])
def test_box_resting_on_ground_has_four_corner_contacts(orc, yaw):
    he = np.array([0.5, 0.4, 0.3])
    pen = 0.01
    q = quat_axis((0, 0, 1), yaw)
    r, _ = manifold(orc, [GROUND, ("cuboid", he)], [GROUND_POSE, pose((0.3, -0.2, he[2] - pen), q)], [(0, 1)])
    c = contacts(r, 0)
    assert r["count"][0] == 4
    corners = np.array([[sx * he[0], sy * he[1], -he[2]] for sx in (-1, 1) for sy in (-1, 1)])
    from tests.test_oracle_queries import qrot
    want = np.array([qrot(q, p) for p in corners]) + np.array([0.3, -0.2, he[2] - pen])
    got = c["w2"].astype(float)
    for p in want:
        assert np.min(np.linalg.norm(got - p, axis=1)) < TOL
    assert np.allclose(c["w1"][:, 2], 0.0, atol=TOL)                  # on the ground's top face
    assert np.allclose(c["w1"][:, :2], c["w2"][:, :2], atol=TOL)
    assert np.allclose(c["depth"], pen, atol=TOL)
    assert np.allclose(c["n"], [0, 0, 1], atol=1e-6)
    assert np.all(c["f1"] == (2 << 30) | 2)                              # the ground's +z face
    assert np.all((c["f2"] >> 30) == 0)                                  # the box's corners


def test_square_on_square_turned_45_degrees_gives_an_octagon(orc):
    he = np.array([0.5, 0.5, 0.3])
    pen = 0.02
    q = quat_axis((0, 0, 1), np.pi / 4)
    r, _ = manifold(orc, [("cuboid", he), ("cuboid", he)], [pose((0, 0, 0)), pose((0, 0, 2 * he[2] - pen), q)],
                    [(0, 1)])
    c = contacts(r, 0)
    assert r["count"][0] == 8
    # the octagon of the two squares: vertices at radius 0.5 / cos(22.5 deg) on the 22.5 + k 45 deg rays
    ang = np.arctan2(c["w1"][:, 1], c["w1"][:, 0])
    rad = np.hypot(c["w1"][:, 0], c["w1"][:, 1])
    assert np.allclose(rad, 0.5 / np.cos(np.pi / 8), atol=1e-4)
    assert np.allclose(np.sort(np.mod(ang - np.pi / 8, np.pi / 4)) % (np.pi / 4), 0, atol=1e-4) or \
        np.allclose(np.abs(np.mod(ang - np.pi / 8 + 1e-3, np.pi / 4) - 1e-3), 0, atol=1e-3)
    assert np.allclose(c["depth"], pen, atol=TOL)
    assert np.all((c["f1"] >> 30) == 1) and np.all((c["f2"] >> 30) == 1)  # edge against edge


def test_ball_on_box_one_contact(orc):
    he = np.array([0.6, 0.6, 0.3])
    rb, pen = 0.5, 0.02
    r, _ = manifold(orc, [("ball", rb), ("cuboid", he)], [pose((0.1, 0.2, he[2] + rb - pen)), pose((0, 0, 0))],
                    [(0, 1), (1, 0)])
    for k, sign in ((0, -1), (1, 1)):
        c = contacts(r, k)
        assert r["count"][k] == 1
        assert abs(c["depth"][0] - pen) < TOL
        assert np.allclose(c["n"][0], [0, 0, sign], atol=1e-6)
    assert r["feature1"][0, 0] == (2 << 30) | 0 and r["feature2"][0, 0] == (2 << 30) | 2
    assert r["feature1"][1, 0] == (2 << 30) | 2 and r["feature2"][1, 0] == (2 << 30) | 0


def test_capsule_lying_on_ground_touches_at_both_ends(orc):
    hh, rc, pen = 0.6, 0.2, 0.01
    q = quat_axis((0, 0, 1), -np.pi / 2)  # the capsule's axis (local y) along world x
    # separated segment: support features within the angular prediction (0.1 rad); with an angular
    # prediction of 0 the rounding of the GJK normal picks a single end point
    r, _ = manifold(orc, [GROUND, ("capsule", hh, rc)], [GROUND_POSE, pose((0.5, 0.5, rc - pen), q)], [(0, 1)],
                    ang1=0.1, ang2=0.1)
    c = contacts(r, 0)
    assert r["count"][0] == 2
    assert np.allclose(c["depth"], pen, atol=TOL)
    assert np.allclose(np.sort(c["w2"][:, 0]), [0.5 - hh, 0.5 + hh], atol=TOL)
    assert np.allclose(c["w2"][:, 2], -pen, atol=TOL)                    # moved out by the radius
    assert sorted(c["f2"].tolist()) == [(2 << 30) | 0, (2 << 30) | 1]    # the end caps


def test_capsule_standing_touches_once_and_prediction_grows_by_radius(orc):
    hh, rc = 0.6, 0.2
    q = quat_axis((1, 0, 0), np.pi / 2)  # axis along world z
    for gap, want in ((-0.01, 1), (0.05, 0)):
        r, _ = manifold(orc, [GROUND, ("capsule", hh, rc)], [GROUND_POSE, pose((0, 0, hh + rc + gap), q)], [(0, 1)],
                        linear=0.0)
        assert r["count"][0] == want
        if want:
            assert abs(contacts(r, 0)["depth"][0] + gap) < TOL
    # with a linear prediction of 0.1 the 0.05 gap is a (negative-depth) contact
    r, _ = manifold(orc, [GROUND, ("capsule", hh, rc)], [GROUND_POSE, pose((0, 0, hh + rc + 0.05), q)], [(0, 1)],
                    linear=0.1)
    assert r["count"][0] == 1 and abs(contacts(r, 0)["depth"][0] + 0.05) < TOL


def test_coincident_clip_points_merge_within_the_tracking_distance(orc):
    he = np.array([0.5, 0.5, 0.3])
    pen = 0.01
    r, store = manifold(orc, [("cuboid", he), ("cuboid", he)], [pose((0, 0, 0)), pose((0, 0, 2 * he[2] - pen))],
                        [(0, 1)])
    assert r["count"][0] == 4
    # the next update, moved by less than the tracking distance: still four, same tracked points
    r2, _ = manifold(orc, [("cuboid", he), ("cuboid", he)], [pose((0, 0, 0)), pose((0.005, 0, 2 * he[2] - pen))],
                     [(0, 1)], store=store)
    assert r2["count"][0] == 4
    assert np.allclose(np.sort(r2["depth"][0, :4]), pen, atol=TOL)


def test_last_gjk_dir_is_kept_between_updates(orc):
    rng = np.random.default_rng(3)
    from tests.test_oracle_queries import rand_hull, rand_quat
    a, b = rand_hull(rng), rand_hull(rng)
    shapes = [("hull", a), ("hull", b)]
    poses = [pose((0, 0, 0), rand_quat(rng)), pose((0.5, 0.3, 0.2), rand_quat(rng))]
    r, store = manifold(orc, shapes, poses, [(0, 1)])
    d1 = r["dirs"][0].copy()
    assert np.linalg.norm(d1) > 0.99
    # the pair's generator keeps the result direction of its last GJK/EPA
    gen = orc.polyhedra_gjk_step(shapeset(orc, shapes, dtype=np.float32), np.stack(poses), [0], [1], 0.0,
                                 np.zeros((1, 3), np.float32))
    assert np.array_equal(d1.view(np.uint32), gen["dir"][0].view(np.uint32))
    poses[1] = pose((0.52, 0.3, 0.2), poses[1][3:])
    r2, _ = manifold(orc, shapes, poses, [(0, 1)], store=store)
    gen2 = orc.polyhedra_gjk_step(shapeset(orc, shapes, dtype=np.float32), np.stack(poses), [0], [1], 0.0, d1[None])
    assert np.array_equal(r2["dirs"][0].view(np.uint32), gen2["dir"][0].view(np.uint32))


def _hull_checks(hp, P):
    n = len(hp)
    nf = len(P["face_first"])
    kept = np.concatenate([P["loop_edge"][P["face_first"][f]:P["face_first"][f] + P["face_count"][f]]
                           for f in range(nf)])
    ne = len(np.unique(kept))                                            # edges between faces
    assert len(kept) == 2 * ne                                           # each bounds two faces
    used = np.count_nonzero(P["vertex_count"])
    assert used + nf - ne == 2                                           # Euler's characteristic
    assert np.allclose(np.linalg.norm(P["face_normals"], axis=1), 1, atol=1e-6)
    assert np.allclose(np.linalg.norm(P["edge_dirs"], axis=1), 1, atol=1e-6)
    for f in range(nf):
        loop = P["loop_vertex"][P["face_first"][f]:P["face_first"][f] + P["face_count"][f]]
        nn = P["face_normals"][f].astype(float)
        d = (hp[loop].astype(float) - hp[loop[0]]) @ nn
        # planar within the merge angle (normals closer than acos(1 - sqrt(eps)) ~ 1.5 degrees merge,
        # so a merged face bends by at most a few hundredths over these unit-sized hulls)
        assert np.all(np.abs(d) < 0.02)
        assert np.all((hp.astype(float) - hp[loop[0]]) @ nn < 0.02)      # every point behind the face
        for v in loop:                                                   # adjacency lists the face
            a = P["adj_face"][P["vertex_first"][v]:P["vertex_first"][v] + P["vertex_count"][v]]
            assert f in a


def test_hull_polyhedra_library_equals_oracle_and_invariants(orc):
    rng = np.random.default_rng(11)
    from ncollide_amd import workloads
    pool, hoff, hcnt = workloads.polyhedral_hulls(rng, 60, 8, 128)
    for o, c in zip(hoff, hcnt):
        hp = pool[o:o + c]
        tr = nh.hull_triangles(hp)
        a = nh.hull_polyhedron(hp, tr)
        b = orc.hull_polyhedron(hp, tr)
        assert a is not None and b is not None
        for k in a:
            x, y = np.asarray(a[k]), np.asarray(b[k])
            assert np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                                  y.view(np.uint32) if y.dtype == np.float32 else y), k
        _hull_checks(hp, a)
    # a box's eight corners: six quadrilateral faces with the axis normals
    c = np.array([[x, y, z] for x in (-1, 1) for y in (-2, 2) for z in (-.5, .5)], np.float32)
    P = nh.hull_polyhedron(c, nh.hull_triangles(c))
    assert sorted(P["face_count"].tolist()) == [4] * 6
    assert sorted(map(tuple, np.abs(P["face_normals"]).round(6).tolist())) == \
        sorted([(1, 0, 0), (1, 0, 0), (0, 1, 0), (0, 1, 0), (0, 0, 1), (0, 0, 1)])
    _hull_checks(c, P)


def test_hull_whose_merged_faces_break_euler_is_rejected_by_both(orc):
    rng = np.random.default_rng(0)
    from ncollide_amd import workloads
    rejected = 0
    for _ in range(200):
        pts, offs, cnts = workloads.ellipsoid_clouds(rng, 1, 8, 200)
        hp = pts[nh.convex_hull_vertices(pts)]
        tr = nh.hull_triangles(hp)
        a, b = nh.hull_polyhedron(hp, tr), orc.hull_polyhedron(hp, tr)
        assert (a is None) == (b is None)
        ok = nh.polyhedra_ok(hp, np.zeros(1, np.uint32), np.array([len(hp)], np.uint32))[0]
        assert ok == (a is not None)
        rejected += a is None
    assert 0 < rejected < 20
