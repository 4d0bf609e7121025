// ncx_oracle.hpp -- CPU restatement of ncollide3d 0.33's collision hot path.
//
// TEST INFRASTRUCTURE ONLY.  This code is the parity checker for the HIP
// library in ncollide_amd/.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product never links it.
//
// Every function restates one reference function operation by operation
// (citations are file:line under the reference crate's src/).  It is
// templated on the scalar type so the same code gives the f32 result the
// GPU must match and the f64 "ncollide3d<f64>" result used for error reports.
// Compile with -ffp-contract=off: the reference (Rust) never contracts a*b+c.
//
// Arithmetic conventions fixed here (nalgebra 0.30 is not vendored, so these
// are the restatement's choices; both the oracle and the GPU follow them):
//   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)    = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   norm(v)       = sqrt(dot(v,v));  normalize(v) = v / norm(v) (component division)
//   Unit::try_new(v, eps): n = norm(v); None if n <= eps else v / n
//   q * v         = t = 2*(q.xyz x v); (t*q.w + q.xyz x t) + v
//   q^-1 * v      = same formula with (-q.xyz, q.w)
//   iso * p       = (q * p) + t ;  iso^-1 * p = q^-1 * (p - t)
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <algorithm>
#include <functional>

namespace ncxo {

template <class R> struct Eps;
template <> struct Eps<float> { static float eps() { return 1.1920929e-07f; } static float maxv() { return 3.40282347e+38f; } };
template <> struct Eps<double> { static double eps() { return 2.2204460492503131e-16; } static double maxv() { return 1.7976931348623157e+308; } };

// ---------------------------------------------------------------- vectors
template <class R> struct V3 {
  R x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return V3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return V3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> inline V3<R> operator-(V3<R> a) { return V3<R>(-a.x, -a.y, -a.z); }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return V3<R>(a.x * s, a.y * s, a.z * s); }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return V3<R>(a.x / s, a.y / s, a.z / s); }
template <class R> inline R dot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> inline R norm_sq(V3<R> a) { return dot(a, a); }
template <class R> inline R norm(V3<R> a) { return std::sqrt(norm_sq(a)); }
template <class R> inline V3<R> normalize(V3<R> a) { return a / norm(a); }
// nalgebra Unit::try_new_and_get: None when norm <= min_norm.
template <class R> inline bool try_normalize(V3<R> v, R min_norm, V3<R>* out, R* n_out = nullptr) {
  R n = norm(v);
  if (n <= min_norm) return false;
  *out = v / n;
  if (n_out) *n_out = n;
  return true;
}
template <class R> inline V3<R> vmin(V3<R> a, V3<R> b) { return V3<R>(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)); }
template <class R> inline V3<R> vmax(V3<R> a, V3<R> b) { return V3<R>(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)); }

template <class R> struct Quat { R i, j, k, w; };
template <class R> inline V3<R> qrot(const Quat<R>& q, V3<R> v) {
  V3<R> qv(q.i, q.j, q.k);
  V3<R> t = cross(qv, v) * R(2);
  V3<R> c = cross(qv, t);
  return (t * q.w + c) + v;
}
template <class R> inline V3<R> qrot_inv(const Quat<R>& q, V3<R> v) {
  Quat<R> c{-q.i, -q.j, -q.k, q.w};
  return qrot(c, v);
}
template <class R> struct Iso {
  V3<R> t;
  Quat<R> q;
  static Iso identity() { Iso m; m.t = V3<R>(0, 0, 0); m.q = Quat<R>{0, 0, 0, 1}; return m; }
};
template <class R> inline V3<R> xform_pt(const Iso<R>& m, V3<R> p) { return qrot(m.q, p) + m.t; }
template <class R> inline V3<R> xform_vec(const Iso<R>& m, V3<R> v) { return qrot(m.q, v); }
template <class R> inline V3<R> inv_xform_pt(const Iso<R>& m, V3<R> p) { return qrot_inv(m.q, p - m.t); }
template <class R> inline V3<R> inv_xform_vec(const Iso<R>& m, V3<R> v) { return qrot_inv(m.q, v); }
// UnitQuaternion::to_rotation_matrix (row-major), used by the cuboid AABB.
template <class R> inline void quat_to_mat(const Quat<R>& q, R m[3][3]) {
  R i = q.i, j = q.j, k = q.k, w = q.w;
  R ww = w * w, ii = i * i, jj = j * j, kk = k * k;
  R ij = i * j * R(2), wk = w * k * R(2), wj = w * j * R(2);
  R ik = i * k * R(2), jk = j * k * R(2), wi = w * i * R(2);
  m[0][0] = ww + ii - jj - kk; m[0][1] = ij - wk; m[0][2] = wj + ik;
  m[1][0] = wk + ij; m[1][1] = ww - ii + jj - kk; m[1][2] = jk - wi;
  m[2][0] = ik - wj; m[2][1] = wi + jk; m[2][2] = ww - ii - jj + kk;
}

template <class R> inline R gjk_eps_tol() { return Eps<R>::eps() * R(10); }  // gjk.rs:26-29

// ---------------------------------------------------------------- shapes
// SEGMENT: a capsule's core segment (Capsule::segment, shape/capsule.rs:54-61), a = (0, -a, 0),
// b = (0, a, 0); the contact-manifold generators use it in place of the capsule.
enum ShapeType : uint32_t { BALL = 0, CUBOID = 1, CAPSULE = 2, HULL = 3, ORIGIN = 4, TRIANGLE = 5, SEGMENT = 6 };
template <class R> struct HullPolyO;
// TRIANGLE: a part of a TriMesh (TriMesh::triangle_at, shape/trimesh.rs:471-475); pts = its three
// vertices a, b, c in the mesh frame.

template <class R> struct Shape {
  uint32_t type;
  R a, b, c;              // ball: r; cuboid: hx,hy,hz; capsule: half_height, radius
  const V3<R>* pts;       // hull vertices (local frame)
  uint32_t npts;
  const HullPolyO<R>* poly = nullptr;  // hull faces, edges and adjacency (contact manifolds only)
};

// utils/point_cloud_support_point.rs:6-24 (strict '>' keeps the first maximum)
template <class R> inline uint32_t point_cloud_support_id(V3<R> dir, const V3<R>* pts, uint32_t n) {
  uint32_t best = 0;
  R best_dot = dot(pts[0], dir);
  for (uint32_t i = 1; i < n; ++i) {
    R d = dot(pts[i], dir);
    if (d > best_dot) { best_dot = d; best = i; }
  }
  return best;
}

// SupportMap::support_point (shape/support_map.rs:26-29) with the overrides of
// Ball (ball.rs:31-38), Cuboid (cuboid.rs:135-146), Capsule (capsule.rs:72-84),
// ConvexHull (convex.rs:418-423) and ConstantOrigin (special_support_maps.rs:9-13).
template <class R> inline V3<R> support_point(const Shape<R>& g, const Iso<R>& m, V3<R> dir) {
  switch (g.type) {
    case BALL: {
      V3<R> u = normalize(dir);
      return m.t + u * g.a;
    }
    case CUBOID: {
      V3<R> ld = inv_xform_vec(m, dir);
      V3<R> p(std::copysign(g.a, ld.x), std::copysign(g.b, ld.y), std::copysign(g.c, ld.z));
      return xform_pt(m, p);
    }
    case CAPSULE: {
      V3<R> ld = inv_xform_vec(m, dir);
      V3<R> u = normalize(ld);
      V3<R> res(R(0), std::copysign(g.a, u.y), R(0));
      V3<R> p = res + u * g.b;
      return xform_pt(m, p);
    }
    case HULL: {
      V3<R> ld = inv_xform_vec(m, dir);
      uint32_t id = point_cloud_support_id(ld, g.pts, g.npts);
      return xform_pt(m, g.pts[id]);
    }
    case TRIANGLE: {  // Triangle::local_support_point (shape/triangle.rs:290-307)
      V3<R> ld = inv_xform_vec(m, dir);
      R d1 = dot(g.pts[0], ld), d2 = dot(g.pts[1], ld), d3 = dot(g.pts[2], ld);
      V3<R> p = d1 > d2 ? (d1 > d3 ? g.pts[0] : g.pts[2]) : (d2 > d3 ? g.pts[1] : g.pts[2]);
      return xform_pt(m, p);
    }
    case SEGMENT: {  // Segment::local_support_point (shape/segment.rs:182-191)
      V3<R> ld = inv_xform_vec(m, dir);
      V3<R> a(R(0), -g.a, R(0)), b(R(0), g.a, R(0));
      return xform_pt(m, dot(a, ld) > dot(b, ld) ? a : b);
    }
    default:
      return V3<R>(0, 0, 0);
  }
}

// ---------------------------------------------------------------- CSO point (query/algorithms/cso_point.rs:13-105)
template <class R> struct CSOPoint {
  V3<R> point, orig1, orig2;
};
template <class R> inline CSOPoint<R> cso_from_shapes(const Iso<R>& m1, const Shape<R>& g1, const Iso<R>& m2,
                                                      const Shape<R>& g2, V3<R> dir) {
  CSOPoint<R> c;
  c.orig1 = support_point(g1, m1, dir);
  c.orig2 = support_point(g2, m2, -dir);
  c.point = c.orig1 - c.orig2;
  return c;
}

// ---------------------------------------------------------------- Voronoi projections
// Location codes shared by the three projections.
enum LocKind { ON_VERTEX = 0, ON_EDGE = 1, ON_FACE = 2, ON_SOLID = 3 };
template <class R> struct Loc {
  int kind;
  int id;
  R bc[3];
};

// query/point/point_segment.rs:52-91 (identity isometry, point = origin)
template <class R> inline V3<R> project_segment(V3<R> a, V3<R> b, V3<R> p, Loc<R>* loc) {
  Iso<R> m = Iso<R>::identity();
  V3<R> ls_pt = inv_xform_pt(m, p);
  V3<R> ab = b - a;
  V3<R> ap = ls_pt - a;
  R ab_ap = dot(ab, ap);
  R sqnab = norm_sq(ab);
  if (ab_ap <= R(0)) {
    loc->kind = ON_VERTEX; loc->id = 0;
    return xform_pt(m, a);
  } else if (ab_ap >= sqnab) {
    loc->kind = ON_VERTEX; loc->id = 1;
    return xform_pt(m, b);
  } else {
    R u = ab_ap / sqnab;
    loc->kind = ON_EDGE; loc->id = 0;
    loc->bc[0] = R(1) - u; loc->bc[1] = u;
    V3<R> proj = a + ab * u;
    return xform_pt(m, proj);
  }
}

// query/point/point_triangle.rs:61-309 (dim3, fixed-point feature off, solid as given)
template <class R> inline V3<R> project_triangle(V3<R> a, V3<R> b, V3<R> c, V3<R> pt, bool solid, Loc<R>* loc) {
  Iso<R> m = Iso<R>::identity();
  V3<R> p = inv_xform_pt(m, pt);
  V3<R> ab = b - a, ac = c - a, ap = p - a;
  R ab_ap = dot(ab, ap), ac_ap = dot(ac, ap);
  if (ab_ap <= R(0) && ac_ap <= R(0)) { loc->kind = ON_VERTEX; loc->id = 0; return xform_pt(m, a); }
  V3<R> bp = p - b;
  R ab_bp = dot(ab, bp), ac_bp = dot(ac, bp);
  if (ab_bp >= R(0) && ac_bp <= ab_bp) { loc->kind = ON_VERTEX; loc->id = 1; return xform_pt(m, b); }
  V3<R> cp = p - c;
  R ab_cp = dot(ab, cp), ac_cp = dot(ac, cp);
  if (ac_cp >= R(0) && ab_cp <= ac_cp) { loc->kind = ON_VERTEX; loc->id = 2; return xform_pt(m, c); }
  V3<R> bc = c - b;
  // stable_check_edges_voronoi (dim3): point_triangle.rs:158-191
  V3<R> n = cross(ab, ac);
  R vc = dot(n, cross(ab, ap));
  if (vc < R(0) && ab_ap >= R(0) && ab_bp <= R(0)) {
    R v = ab_ap / norm_sq(ab);
    loc->kind = ON_EDGE; loc->id = 0; loc->bc[0] = R(1) - v; loc->bc[1] = v;
    return xform_pt(m, a + ab * v);
  }
  R vb = -dot(n, cross(ac, cp));
  if (vb < R(0) && ac_ap >= R(0) && ac_cp <= R(0)) {
    R w = ac_ap / norm_sq(ac);
    loc->kind = ON_EDGE; loc->id = 2; loc->bc[0] = R(1) - w; loc->bc[1] = w;
    return xform_pt(m, a + ac * w);
  }
  R va = dot(n, cross(bc, bp));
  if (va < R(0) && ac_bp - ab_bp >= R(0) && ab_cp - ac_cp >= R(0)) {
    R w = dot(bc, bp) / norm_sq(bc);
    loc->kind = ON_EDGE; loc->id = 1; loc->bc[0] = R(1) - w; loc->bc[1] = w;
    return xform_pt(m, b + bc * w);
  }
  int clockwise = dot(n, ap) >= R(0) ? 0 : 1;
  if (va + vb + vc != R(0)) {
    R denom = R(1) / (va + vb + vc);
    R v = vb * denom;
    R w = vc * denom;
    loc->kind = ON_FACE; loc->id = clockwise;
    loc->bc[0] = R(1) - v - w; loc->bc[1] = v; loc->bc[2] = w;
    V3<R> res = a + ab * v + ac * w;
    return xform_pt(m, res);
  }
  if (solid) { loc->kind = ON_SOLID; loc->id = 0; return pt; }
  // Non-solid fallback (point_triangle.rs:261-307); unused by GJK/EPA (solid=true).
  R v = ab_ap / (ab_ap - ab_bp);
  R w = ac_ap / (ac_ap - ac_cp);
  R u = (ac_bp - ab_bp) / (ac_bp - ab_bp + ab_cp - ac_cp);
  R d_ab = norm_sq(ap) - (norm_sq(ab) * v * v);
  R d_ac = norm_sq(ap) - (norm_sq(ac) * u * u);
  R d_bc = norm_sq(bp) - (norm_sq(bc) * w * w);
  if (d_ab < d_ac) {
    if (d_ab < d_bc) { loc->kind = ON_EDGE; loc->id = 0; loc->bc[0] = R(1) - v; loc->bc[1] = v; return xform_pt(m, a + ab * v); }
    loc->kind = ON_EDGE; loc->id = 1; loc->bc[0] = R(1) - u; loc->bc[1] = u; return xform_pt(m, b + bc * u);
  }
  if (d_ac < d_bc) { loc->kind = ON_EDGE; loc->id = 2; loc->bc[0] = R(1) - w; loc->bc[1] = w; return xform_pt(m, a + ac * w); }
  loc->kind = ON_EDGE; loc->id = 1; loc->bc[0] = R(1) - u; loc->bc[1] = u; return xform_pt(m, b + bc * u);
}

// query/point/point_tetrahedron.rs:35-353 (solid = true)
template <class R> inline V3<R> project_tetrahedron(V3<R> A, V3<R> B, V3<R> C, V3<R> D, V3<R> pt, Loc<R>* loc) {
  Iso<R> m = Iso<R>::identity();
  V3<R> p = inv_xform_pt(m, pt);
  V3<R> ab = B - A, ac = C - A, ad = D - A, ap = p - A;
  R ap_ab = dot(ap, ab), ap_ac = dot(ap, ac), ap_ad = dot(ap, ad);
  if (ap_ab <= R(0) && ap_ac <= R(0) && ap_ad <= R(0)) { loc->kind = ON_VERTEX; loc->id = 0; return xform_pt(m, A); }
  V3<R> bc = C - B, bd = D - B, bp = p - B;
  R bp_bc = dot(bp, bc), bp_bd = dot(bp, bd), bp_ab = dot(bp, ab);
  if (bp_bc <= R(0) && bp_bd <= R(0) && bp_ab >= R(0)) { loc->kind = ON_VERTEX; loc->id = 1; return xform_pt(m, B); }
  V3<R> cd = D - C, cp = p - C;
  R cp_ac = dot(cp, ac), cp_bc = dot(cp, bc), cp_cd = dot(cp, cd);
  if (cp_cd <= R(0) && cp_bc >= R(0) && cp_ac >= R(0)) { loc->kind = ON_VERTEX; loc->id = 2; return xform_pt(m, C); }
  V3<R> dp = p - D;
  R dp_cd = dot(dp, cd), dp_bd = dot(dp, bd), dp_ad = dot(dp, ad);
  if (dp_ad >= R(0) && dp_bd >= R(0) && dp_cd >= R(0)) { loc->kind = ON_VERTEX; loc->id = 3; return xform_pt(m, D); }

  // check_edge (point_tetrahedron.rs:106-152)
  auto check_edge = [&](int i, V3<R> a, V3<R> nabc, V3<R> nabd, V3<R> ap_, V3<R> ab_, R ap_ab_, R bp_ab_,
                        R* dabc, R* dabd, V3<R>* out) -> bool {
    R ab_ab = ap_ab_ - bp_ab_;
    V3<R> ap_x_ab = cross(ap_, ab_);
    *dabc = dot(ap_x_ab, nabc);
    *dabd = dot(ap_x_ab, nabd);
    if (ab_ab != R(0) && *dabc >= R(0) && *dabd >= R(0) && ap_ab_ >= R(0) && ap_ab_ <= ab_ab) {
      R u = ap_ab_ / ab_ab;
      loc->kind = ON_EDGE; loc->id = i; loc->bc[0] = R(1) - u; loc->bc[1] = u;
      *out = xform_pt(m, a + ab_ * u);
      return true;
    }
    return false;
  };
  V3<R> out;
  V3<R> nabc = cross(ab, ac), nabd = cross(ab, ad);
  R dabc, dabd, dacd, dacb, dadb, dadc, dbca, dbcd, dbdc, dbda, dcda, dcdb;
  if (check_edge(0, A, nabc, nabd, ap, ab, ap_ab, bp_ab, &dabc, &dabd, &out)) return out;
  V3<R> nacd = cross(ac, ad);
  if (check_edge(1, A, nacd, -nabc, ap, ac, ap_ac, cp_ac, &dacd, &dacb, &out)) return out;
  if (check_edge(2, A, -nabd, -nacd, ap, ad, ap_ad, dp_ad, &dadb, &dadc, &out)) return out;
  V3<R> nbcd = cross(bc, bd);
  if (check_edge(3, B, nabc, nbcd, bp, bc, bp_bc, cp_bc, &dbca, &dbcd, &out)) return out;
  if (check_edge(4, B, -nbcd, nabd, bp, bd, bp_bd, dp_bd, &dbdc, &dbda, &out)) return out;
  if (check_edge(5, C, nacd, nbcd, cp, cd, cp_cd, dp_cd, &dcda, &dcdb, &out)) return out;

  // check_face (point_tetrahedron.rs:251-306)
  auto check_face = [&](int i, V3<R> a, V3<R> b, V3<R> c, V3<R> ap_, V3<R> bp_, V3<R> cp_, V3<R> ab_, V3<R> ac_,
                        V3<R> ad_, R d1, R d2, R d3, V3<R>* o) -> bool {
    if (d1 < R(0) && d2 < R(0) && d3 < R(0)) {
      V3<R> n = cross(ab_, ac_);
      if (dot(n, ad_) * dot(n, ap_) < R(0)) {
        V3<R> normal;
        if (!try_normalize(n, Eps<R>::eps(), &normal)) return false;
        R vc = dot(normal, cross(ap_, bp_));
        R va = dot(normal, cross(bp_, cp_));
        R vb = dot(normal, cross(cp_, ap_));
        R denom = va + vb + vc;
        R inv_denom = R(1) / denom;
        R bcs[3] = {va * inv_denom, vb * inv_denom, vc * inv_denom};
        V3<R> res = a * bcs[0] + b * bcs[1] + c * bcs[2];
        loc->kind = ON_FACE; loc->id = i; loc->bc[0] = bcs[0]; loc->bc[1] = bcs[1]; loc->bc[2] = bcs[2];
        *o = xform_pt(m, res);
        return true;
      }
    }
    return false;
  };
  if (check_face(0, A, B, C, ap, bp, cp, ab, ac, ad, dabc, dbca, dacb, &out)) return out;
  if (check_face(1, A, B, D, ap, bp, dp, ab, ad, ac, dadb, dabd, dbda, &out)) return out;
  if (check_face(2, A, C, D, ap, cp, dp, ac, ad, ab, dacd, dcda, dadc, &out)) return out;
  if (check_face(3, B, C, D, bp, cp, dp, bc, bd, -ab, dbcd, dcdb, dbdc, &out)) return out;
  loc->kind = ON_SOLID; loc->id = 0;
  return xform_pt(m, p);
}

// ---------------------------------------------------------------- VoronoiSimplex (voronoi_simplex3.rs:12-352)
template <class R> struct VoronoiSimplex {
  int prev_vertices[4] = {0, 1, 2, 3};
  R prev_proj[3] = {0, 0, 0};
  int prev_dim = 0;
  CSOPoint<R> vertices[4];
  R proj[3] = {0, 0, 0};
  int dim = 0;

  void swap(int i1, int i2) {
    std::swap(vertices[i1], vertices[i2]);
    std::swap(prev_vertices[i1], prev_vertices[i2]);
  }
  void reset(const CSOPoint<R>& pt) { dim = 0; prev_dim = 0; vertices[0] = pt; }
  bool add_point(const CSOPoint<R>& pt) {
    prev_dim = dim;
    prev_proj[0] = proj[0]; prev_proj[1] = proj[1]; prev_proj[2] = proj[2];
    prev_vertices[0] = 0; prev_vertices[1] = 1; prev_vertices[2] = 2; prev_vertices[3] = 3;
    R tol = gjk_eps_tol<R>();
    if (dim == 0) {
      if (norm_sq(vertices[0].point - pt.point) < tol) return false;
    } else if (dim == 1) {
      V3<R> ab = vertices[1].point - vertices[0].point;
      V3<R> ac = pt.point - vertices[0].point;
      if (norm_sq(cross(ab, ac)) < tol) return false;
    } else {
      V3<R> ab = vertices[1].point - vertices[0].point;
      V3<R> ac = vertices[2].point - vertices[0].point;
      V3<R> ap = pt.point - vertices[0].point;
      V3<R> n = normalize(cross(ab, ac));
      if (std::fabs(dot(n, ap)) < tol) return false;
    }
    dim += 1;
    vertices[dim] = pt;
    return true;
  }
  V3<R> project_origin_and_reduce() {
    V3<R> o(0, 0, 0);
    if (dim == 0) {
      proj[0] = R(1);
      return vertices[0].point;
    } else if (dim == 1) {
      Loc<R> loc;
      V3<R> pr = project_segment(vertices[0].point, vertices[1].point, o, &loc);
      if (loc.kind == ON_VERTEX && loc.id == 0) { proj[0] = R(1); dim = 0; }
      else if (loc.kind == ON_VERTEX) { swap(0, 1); proj[0] = R(1); dim = 0; }
      else { proj[0] = loc.bc[0]; proj[1] = loc.bc[1]; }
      return pr;
    } else if (dim == 2) {
      Loc<R> loc;
      V3<R> pr = project_triangle(vertices[0].point, vertices[1].point, vertices[2].point, o, true, &loc);
      if (loc.kind == ON_VERTEX) { swap(0, loc.id); proj[0] = R(1); dim = 0; }
      else if (loc.kind == ON_EDGE) {
        if (loc.id == 0) { proj[0] = loc.bc[0]; proj[1] = loc.bc[1]; }
        else if (loc.id == 1) { swap(0, 2); proj[0] = loc.bc[1]; proj[1] = loc.bc[0]; }
        else { swap(1, 2); proj[0] = loc.bc[0]; proj[1] = loc.bc[1]; }
        dim = 1;
      } else if (loc.kind == ON_FACE) {
        proj[0] = loc.bc[0]; proj[1] = loc.bc[1]; proj[2] = loc.bc[2];
      }
      return pr;
    } else {
      Loc<R> loc;
      V3<R> pr = project_tetrahedron(vertices[0].point, vertices[1].point, vertices[2].point, vertices[3].point, o, &loc);
      if (loc.kind == ON_VERTEX) { swap(0, loc.id); proj[0] = R(1); dim = 0; }
      else if (loc.kind == ON_EDGE) {
        switch (loc.id) {
          case 0: break;
          case 1: swap(1, 2); break;
          case 2: swap(1, 3); break;
          case 3: swap(0, 2); break;
          case 4: swap(0, 3); break;
          case 5: swap(0, 2); swap(1, 3); break;
        }
        if (loc.id == 3 || loc.id == 4) { proj[0] = loc.bc[1]; proj[1] = loc.bc[0]; }
        else { proj[0] = loc.bc[0]; proj[1] = loc.bc[1]; }
        dim = 1;
      } else if (loc.kind == ON_FACE) {
        switch (loc.id) {
          case 0: proj[0] = loc.bc[0]; proj[1] = loc.bc[1]; proj[2] = loc.bc[2]; break;
          case 1: vertices[2] = vertices[3]; proj[0] = loc.bc[0]; proj[1] = loc.bc[1]; proj[2] = loc.bc[2]; break;
          case 2: vertices[1] = vertices[3]; proj[0] = loc.bc[0]; proj[1] = loc.bc[2]; proj[2] = loc.bc[1]; break;
          case 3: vertices[0] = vertices[3]; proj[0] = loc.bc[2]; proj[1] = loc.bc[0]; proj[2] = loc.bc[1]; break;
        }
        dim = 2;
      }
      return pr;
    }
  }
  const CSOPoint<R>& prev_point(int i) const { return vertices[prev_vertices[i]]; }
};

// gjk.rs:367-388
template <class R> inline void gjk_result(const VoronoiSimplex<R>& s, bool prev, V3<R>* p1, V3<R>* p2) {
  V3<R> r1(0, 0, 0), r2(0, 0, 0);
  if (prev) {
    for (int i = 0; i < s.prev_dim + 1; ++i) {
      R coord = s.prev_proj[i];
      const CSOPoint<R>& pt = s.prev_point(i);
      r1 = r1 + pt.orig1 * coord;
      r2 = r2 + pt.orig2 * coord;
    }
  } else {
    for (int i = 0; i < s.dim + 1; ++i) {
      R coord = s.proj[i];
      const CSOPoint<R>& pt = s.vertices[i];
      r1 = r1 + pt.orig1 * coord;
      r2 = r2 + pt.orig2 * coord;
    }
  }
  *p1 = r1; *p2 = r2;
}

// ---------------------------------------------------------------- GJK (gjk.rs:76-177)
// GJK_CAP: the loop stopped at a caller-set iteration cap below the reference's 10000
// (the reference's own cap returns NoIntersection(+x), gjk.rs:172-175).
enum GJKStatus { GJK_INTERSECTION = 0, GJK_CLOSEST = 1, GJK_PROXIMITY = 2, GJK_NO_INTERSECTION = 3, GJK_NAN = 4,
                 GJK_CAP = 5 };
template <class R> struct GJKResult {
  int status;
  V3<R> p1, p2, dir;
  int iters;  // support-point evaluations inside the loop (the reference's niter plus the final one)
};

template <class R> inline GJKResult<R> gjk_closest_points(const Iso<R>& m1, const Shape<R>& g1, const Iso<R>& m2,
                                                          const Shape<R>& g2, R max_dist, bool exact_dist,
                                                          VoronoiSimplex<R>& simplex, int max_iter = 10000) {
  GJKResult<R> res;
  res.iters = 0;
  R eps_tol = gjk_eps_tol<R>();
  R eps_rel = std::sqrt(eps_tol);
  V3<R> proj = simplex.project_origin_and_reduce();
  V3<R> old_dir;
  {
    V3<R> pd;
    if (try_normalize(proj, R(0), &pd)) old_dir = -pd;
    else { res.status = GJK_INTERSECTION; return res; }
  }
  R max_bound = Eps<R>::maxv();
  V3<R> dir;
  int niter = 0;
  for (;;) {
    R old_max_bound = max_bound;
    {
      V3<R> nd; R dist;
      if (try_normalize(-proj, eps_tol, &nd, &dist)) { dir = nd; max_bound = dist; }
      else { res.status = GJK_INTERSECTION; return res; }
    }
    if (max_bound >= old_max_bound) {
      if (exact_dist) { gjk_result(simplex, true, &res.p1, &res.p2); res.dir = old_dir; res.status = GJK_CLOSEST; }
      else { res.dir = old_dir; res.status = GJK_PROXIMITY; }
      return res;
    }
    CSOPoint<R> cso = cso_from_shapes(m1, g1, m2, g2, dir);
    res.iters += 1;
    R min_bound = -dot(dir, cso.point);
    if (!(min_bound == min_bound)) { res.status = GJK_NAN; return res; }
    if (min_bound > max_dist) { res.dir = dir; res.status = GJK_NO_INTERSECTION; return res; }
    else if (!exact_dist && min_bound > R(0) && max_bound <= max_dist) { res.dir = old_dir; res.status = GJK_PROXIMITY; return res; }
    else if (max_bound - min_bound <= eps_rel * max_bound) {
      if (exact_dist) { gjk_result(simplex, false, &res.p1, &res.p2); res.dir = dir; res.status = GJK_CLOSEST; }
      else { res.dir = dir; res.status = GJK_PROXIMITY; }
      return res;
    }
    if (!simplex.add_point(cso)) {
      if (exact_dist) { gjk_result(simplex, false, &res.p1, &res.p2); res.dir = dir; res.status = GJK_CLOSEST; }
      else { res.dir = dir; res.status = GJK_PROXIMITY; }
      return res;
    }
    old_dir = dir;
    proj = simplex.project_origin_and_reduce();
    if (simplex.dim == 3) {
      if (min_bound >= eps_tol) {
        if (exact_dist) { gjk_result(simplex, true, &res.p1, &res.p2); res.dir = old_dir; res.status = GJK_CLOSEST; }
        else { res.dir = old_dir; res.status = GJK_PROXIMITY; }
        return res;
      } else { res.status = GJK_INTERSECTION; return res; }
    }
    niter += 1;
    if (niter == max_iter) {
      res.dir = V3<R>(1, 0, 0);
      res.status = max_iter < 10000 ? GJK_CAP : GJK_NO_INTERSECTION;
      return res;
    }
  }
}

// ---------------------------------------------------------------- Rust std BinaryHeap (max-heap) restatement
// push = sift_up(0, len); pop = swap-with-last, sift_down_to_bottom(0) then sift_up.
// `le(a, b)` is PartialOrd::le of the element type.
template <class T, class LE> struct RustHeap {
  std::vector<T> data;
  LE le;
  void sift_up(size_t start, size_t pos) {
    T elt = data[pos];
    while (pos > start) {
      size_t parent = (pos - 1) / 2;
      if (le(elt, data[parent])) break;
      data[pos] = data[parent];
      pos = parent;
    }
    data[pos] = elt;
  }
  void push(const T& v) { data.push_back(v); sift_up(0, data.size() - 1); }
  bool empty() const { return data.empty(); }
  const T& peek() const { return data[0]; }
  bool pop(T* out) {
    if (data.empty()) return false;
    T item = data.back();
    data.pop_back();
    if (!data.empty()) {
      std::swap(item, data[0]);
      // sift_down_to_bottom(0)
      size_t end = data.size();
      size_t pos = 0;
      T elt = data[0];
      size_t child = 1;
      while (end >= 2 && child <= end - 2) {
        child += le(data[child], data[child + 1]) ? 1 : 0;
        data[pos] = data[child];
        pos = child;
        child = 2 * pos + 1;
      }
      if (child == end - 1) { data[pos] = data[child]; pos = child; }
      data[pos] = elt;
      sift_up(0, pos);
    }
    *out = item;
    return true;
  }
};

// ---------------------------------------------------------------- EPA (epa3.rs:171-454)
template <class R> struct EPAFaceId { uint32_t id; R neg_dist; };
template <class R> struct EPAFaceLE { bool operator()(const EPAFaceId<R>& a, const EPAFaceId<R>& b) const { return a.neg_dist <= b.neg_dist; } };

template <class R> struct EPAFace {
  uint32_t pts[3], adj[3];
  V3<R> normal;
  R bcoords[3];
  bool deleted;
};

// utils/triangle.rs:85-102 (relative_eq!(|p1p2 x p1p3|^2, 0, epsilon = (100 eps)^2))
template <class R> inline bool is_affinely_dependent_triangle(V3<R> p1, V3<R> p2, V3<R> p3) {
  V3<R> p1p2 = p2 - p1, p1p3 = p3 - p1;
  R e = Eps<R>::eps() * R(100);
  R x = norm_sq(cross(p1p2, p1p3));
  R epsilon = e * e;
  if (x == R(0)) return true;
  if (std::isinf(x)) return false;
  R diff = std::fabs(x - R(0));
  if (diff <= epsilon) return true;
  R largest = std::fabs(x);
  return diff <= largest * Eps<R>::eps();
}

// The first direction Vector::orthonormal_subspace_basis(&[v], f) hands to its callback for one
// 3D vector (nalgebra): a = (v.z, 0, -v.x) if |v.x| > |v.y| else (0, -v.z, v.y), a /= |a|, and the
// callback receives a x v; epa3.rs:299-305 stops after that first one.  v is not normalised by the
// caller, so the direction has length |v|: it is orthogonal to v and to a.
template <class R> inline V3<R> orthonormal_subspace_first(V3<R> v) {
  V3<R> a;
  if (std::fabs(v.x) > std::fabs(v.y)) a = V3<R>(v.z, R(0), -v.x);
  else a = V3<R>(R(0), -v.z, v.y);
  a = normalize(a);
  return cross(a, v);
}

template <class R> struct EPA {
  std::vector<CSOPoint<R>> vertices;
  std::vector<EPAFace<R>> faces;
  struct Sil { uint32_t face_id, opp_pt_id; };
  std::vector<Sil> silhouette;
  RustHeap<EPAFaceId<R>, EPAFaceLE<R>> heap;
  int iters = 0;

  // Face::new (epa3.rs:96-115) + new_with_proj (:63-94)
  bool make_face(const uint32_t pts[3], const uint32_t adj[3], EPAFace<R>* f) {
    Loc<R> loc;
    project_triangle(vertices[pts[0]].point, vertices[pts[1]].point, vertices[pts[2]].point, V3<R>(0, 0, 0), true, &loc);
    for (int i = 0; i < 3; ++i) { f->pts[i] = pts[i]; f->adj[i] = adj[i]; }
    f->deleted = false;
    // utils/ccw_face_normal.rs:21-27
    V3<R> ab = vertices[pts[1]].point - vertices[pts[0]].point;
    V3<R> ac = vertices[pts[2]].point - vertices[pts[0]].point;
    V3<R> n;
    if (try_normalize(cross(ab, ac), Eps<R>::eps(), &n)) f->normal = n;
    else f->normal = V3<R>(0, 0, 0);
    if (loc.kind == ON_FACE) {
      f->bcoords[0] = loc.bc[0]; f->bcoords[1] = loc.bc[1]; f->bcoords[2] = loc.bc[2];
      return true;
    }
    f->bcoords[0] = f->bcoords[1] = f->bcoords[2] = R(0);
    return false;
  }
  void face_closest_points(const EPAFace<R>& f, V3<R>* p1, V3<R>* p2) {
    *p1 = vertices[f.pts[0]].orig1 * f.bcoords[0] + vertices[f.pts[1]].orig1 * f.bcoords[1] + vertices[f.pts[2]].orig1 * f.bcoords[2];
    *p2 = vertices[f.pts[0]].orig2 * f.bcoords[0] + vertices[f.pts[1]].orig2 * f.bcoords[1] + vertices[f.pts[2]].orig2 * f.bcoords[2];
  }
  static uint32_t next_ccw_pt_id(const EPAFace<R>& f, uint32_t id) {
    if (f.pts[0] == id) return 1;
    if (f.pts[1] == id) return 2;
    return 0;
  }
  bool can_be_seen_by(const EPAFace<R>& f, uint32_t point, uint32_t opp_pt_id) {
    V3<R> p0 = vertices[f.pts[opp_pt_id]].point;
    V3<R> p1 = vertices[f.pts[(opp_pt_id + 1) % 3]].point;
    V3<R> p2 = vertices[f.pts[(opp_pt_id + 2) % 3]].point;
    V3<R> pt = vertices[point].point;
    return dot(pt - p0, f.normal) >= -gjk_eps_tol<R>() || is_affinely_dependent_triangle(p1, p2, pt);
  }
  void compute_silhouette(uint32_t point, uint32_t id, uint32_t opp_pt_id) {
    if (!faces[id].deleted) {
      if (!can_be_seen_by(faces[id], point, opp_pt_id)) {
        silhouette.push_back(Sil{id, opp_pt_id});
      } else {
        faces[id].deleted = true;
        uint32_t adj_pt_id1 = (opp_pt_id + 2) % 3;
        uint32_t adj_pt_id2 = opp_pt_id;
        uint32_t adj1 = faces[id].adj[adj_pt_id1];
        uint32_t adj2 = faces[id].adj[adj_pt_id2];
        uint32_t o1 = next_ccw_pt_id(faces[adj1], faces[id].pts[adj_pt_id1]);
        uint32_t o2 = next_ccw_pt_id(faces[adj2], faces[id].pts[adj_pt_id2]);
        compute_silhouette(point, adj1, o1);
        compute_silhouette(point, adj2, o2);
      }
    }
  }

  // Returns false for the reference's `None`.
  bool closest_points(const Iso<R>& m1, const Shape<R>& g1, const Iso<R>& m2, const Shape<R>& g2,
                      const VoronoiSimplex<R>& simplex, V3<R>* p1, V3<R>* p2, V3<R>* n, int max_iter = 10000) {
    R eps_tol = Eps<R>::eps() * R(100);
    vertices.clear(); faces.clear(); heap.data.clear(); silhouette.clear();
    iters = 0;
    for (int i = 0; i < simplex.dim + 1; ++i) vertices.push_back(simplex.vertices[i]);
    if (simplex.dim == 0) {
      *p1 = V3<R>(0, 0, 0); *p2 = V3<R>(0, 0, 0); *n = V3<R>(0, 1, 0);
      return true;
    } else if (simplex.dim == 3) {
      V3<R> dp1 = vertices[1].point - vertices[0].point;
      V3<R> dp2 = vertices[2].point - vertices[0].point;
      V3<R> dp3 = vertices[3].point - vertices[0].point;
      if (dot(cross(dp1, dp2), dp3) > R(0)) std::swap(vertices[1], vertices[2]);
      const uint32_t P[4][3] = {{0, 1, 2}, {1, 3, 2}, {0, 2, 3}, {0, 3, 1}};
      const uint32_t A[4][3] = {{3, 1, 2}, {3, 2, 0}, {0, 1, 3}, {2, 1, 0}};
      bool inside[4];
      for (int f = 0; f < 4; ++f) { EPAFace<R> face; inside[f] = make_face(P[f], A[f], &face); faces.push_back(face); }
      for (int f = 0; f < 4; ++f) {
        if (inside[f]) {
          R dist = dot(faces[f].normal, vertices[f].point);
          if (-dist > gjk_eps_tol<R>()) return false;  // FaceId::new(..)? -> None
          heap.push(EPAFaceId<R>{(uint32_t)f, -dist});
        }
      }
    } else {
      if (simplex.dim == 1) {
        V3<R> dpt = vertices[1].point - vertices[0].point;
        vertices.push_back(cso_from_shapes(m1, g1, m2, g2, orthonormal_subspace_first(dpt)));
      }
      const uint32_t P1[3] = {0, 1, 2}, P2[3] = {0, 2, 1};
      const uint32_t A1[3] = {1, 1, 1}, A2[3] = {0, 0, 0};
      EPAFace<R> f1, f2;
      make_face(P1, A1, &f1); make_face(P2, A2, &f2);
      faces.push_back(f1); faces.push_back(f2);
      heap.push(EPAFaceId<R>{0, R(0)});
      heap.push(EPAFaceId<R>{1, R(0)});
    }
    int niter = 0;
    R max_dist = Eps<R>::maxv();
    if (heap.empty()) return false;  // the reference's heap.peek().unwrap() would panic here
    EPAFaceId<R> best_face_id = heap.peek();
    EPAFaceId<R> face_id;
    while (heap.pop(&face_id)) {
      EPAFace<R> face = faces[face_id.id];
      if (face.deleted) continue;
      CSOPoint<R> cso = cso_from_shapes(m1, g1, m2, g2, face.normal);
      iters += 1;
      uint32_t support_point_id = (uint32_t)vertices.size();
      vertices.push_back(cso);
      R candidate_max_dist = dot(cso.point, face.normal);
      if (candidate_max_dist < max_dist) { best_face_id = face_id; max_dist = candidate_max_dist; }
      R curr_dist = -face_id.neg_dist;
      if (max_dist - curr_dist < eps_tol) {
        const EPAFace<R>& best = faces[best_face_id.id];
        face_closest_points(best, p1, p2);
        *n = best.normal;
        return true;
      }
      faces[face_id.id].deleted = true;
      uint32_t o1 = next_ccw_pt_id(faces[face.adj[0]], face.pts[0]);
      uint32_t o2 = next_ccw_pt_id(faces[face.adj[1]], face.pts[1]);
      uint32_t o3 = next_ccw_pt_id(faces[face.adj[2]], face.pts[2]);
      compute_silhouette(support_point_id, face.adj[0], o1);
      compute_silhouette(support_point_id, face.adj[1], o2);
      compute_silhouette(support_point_id, face.adj[2], o3);
      uint32_t first_new_face_id = (uint32_t)faces.size();
      if (silhouette.empty()) return false;
      for (size_t s = 0; s < silhouette.size(); ++s) {
        Sil edge = silhouette[s];
        if (!faces[edge.face_id].deleted) {
          uint32_t new_face_id = (uint32_t)faces.size();
          uint32_t pt_id1 = faces[edge.face_id].pts[(edge.opp_pt_id + 2) % 3];
          uint32_t pt_id2 = faces[edge.face_id].pts[(edge.opp_pt_id + 1) % 3];
          uint32_t pts[3] = {pt_id1, pt_id2, support_point_id};
          uint32_t adj[3] = {edge.face_id, new_face_id + 1, new_face_id - 1};
          EPAFace<R> nf;
          bool inside = make_face(pts, adj, &nf);
          faces[edge.face_id].adj[(edge.opp_pt_id + 1) % 3] = new_face_id;
          faces.push_back(nf);
          if (inside) {
            V3<R> pt = vertices[faces[new_face_id].pts[0]].point;
            R dist = dot(faces[new_face_id].normal, pt);
            if (dist < curr_dist) {
              face_closest_points(face, p1, p2);
              *n = face.normal;
              return true;
            }
            if (-dist > gjk_eps_tol<R>()) return false;
            heap.push(EPAFaceId<R>{new_face_id, -dist});
          }
        }
      }
      if (first_new_face_id == faces.size()) return false;
      faces[first_new_face_id].adj[2] = (uint32_t)faces.size() - 1;
      faces.back().adj[1] = first_new_face_id;
      silhouette.clear();
      niter += 1;
      if (niter > max_iter) return false;
    }
    const EPAFace<R>& best = faces[best_face_id.id];
    face_closest_points(best, p1, p2);
    *n = best.normal;
    return true;
  }
};

// ---------------------------------------------------------------- contact / distance dispatch
template <class R> struct Contact {
  V3<R> world1, world2, normal;
  R depth;
};

// contact_support_map_support_map_with_params (contact_support_map_support_map.rs:38-88).
// init_dir == nullptr is the reference's None.  Returns false for `NoIntersection`.
template <class R> inline bool contact_support_map_support_map(const Iso<R>& m1, const Shape<R>& g1, const Iso<R>& m2,
                                                               const Shape<R>& g2, R prediction, const V3<R>* init_dir,
                                                               Contact<R>* out, int* gjk_iters = nullptr,
                                                               int* epa_iters = nullptr, int gjk_max_iter = 10000,
                                                               int* status_out = nullptr) {
  V3<R> dir;
  if (init_dir) dir = *init_dir;
  else if (!try_normalize(m2.t - m1.t, Eps<R>::eps(), &dir)) dir = V3<R>(1, 0, 0);
  VoronoiSimplex<R> simplex;
  simplex.reset(cso_from_shapes(m1, g1, m2, g2, dir));
  GJKResult<R> r = gjk_closest_points(m1, g1, m2, g2, prediction, true, simplex, gjk_max_iter);
  if (gjk_iters) *gjk_iters = r.iters;
  if (epa_iters) *epa_iters = 0;
  if (status_out) *status_out = r.status;
  V3<R> w1, w2, n;
  if (r.status == GJK_CLOSEST) { w1 = r.p1; w2 = r.p2; n = r.dir; }
  else if (r.status == GJK_INTERSECTION) {
    EPA<R> epa;
    bool ok = epa.closest_points(m1, g1, m2, g2, simplex, &w1, &w2, &n);
    if (epa_iters) *epa_iters = epa.iters;
    if (!ok) return false;
  } else {
    return false;
  }
  // Contact::new_wo_depth (contact.rs:43-46)
  out->world1 = w1; out->world2 = w2; out->normal = n;
  out->depth = -dot(n, w2 - w1);
  return true;
}

// contact_ball_ball (contact_ball_ball.rs:8-38)
template <class R> inline bool contact_ball_ball(V3<R> c1, R r1, V3<R> c2, R r2, R prediction, Contact<R>* out) {
  V3<R> delta = c2 - c1;
  R d2 = norm_sq(delta);
  R sum_radius = r1 + r2;
  R sre = sum_radius + prediction;
  if (d2 < sre * sre) {
    V3<R> normal = (d2 != R(0)) ? normalize(delta) : V3<R>(1, 0, 0);
    out->world1 = c1 + normal * r1;
    out->world2 = c2 + normal * (-r2);
    out->normal = normal;
    out->depth = sum_radius - std::sqrt(d2);
    return true;
  }
  return false;
}

// point_projection_on_support_map (point_support_map.rs:15-53), solid = false.
template <class R> inline V3<R> project_point_support_map(const Iso<R>& m_in, const Shape<R>& g, V3<R> point, bool* inside) {
  Iso<R> id = Iso<R>::identity();
  Iso<R> m = m_in;
  m.t = (-point) + m_in.t;
  V3<R> dir;
  if (!try_normalize(-m.t, Eps<R>::eps(), &dir)) dir = V3<R>(1, 0, 0);
  Shape<R> origin{ORIGIN, 0, 0, 0, nullptr, 0};
  VoronoiSimplex<R> simplex;
  simplex.reset(cso_from_shapes(m, g, id, origin, dir));
  GJKResult<R> r = gjk_closest_points(m, g, id, origin, Eps<R>::maxv(), true, simplex);
  if (r.status == GJK_CLOSEST) { *inside = false; return r.p1 + point; }
  // GJKResult::Intersection -> EPA
  EPA<R> epa;
  V3<R> p1, p2, n;
  if (r.status == GJK_INTERSECTION && epa.closest_points(m, g, id, origin, simplex, &p1, &p2, &n)) {
    *inside = true;
    return p1 + point;
  }
  *inside = true;
  return point;
}

// AABB::local_point_projection + project_point_with_feature (point_aabb.rs:9-135), cuboid h, solid=false.
// feature: kind 0 face(id), 1 edge(id), 2 vertex(id), 3 unknown.
template <class R> inline V3<R> project_point_cuboid(const Iso<R>& m, V3<R> h, V3<R> pt, bool* inside_out, int* fkind, int* fid) {
  V3<R> mins = -h, maxs = h;  // Point::origin() + (-h)
  mins = V3<R>(R(0) + mins.x, R(0) + mins.y, R(0) + mins.z);
  maxs = V3<R>(R(0) + maxs.x, R(0) + maxs.y, R(0) + maxs.z);
  V3<R> ls_pt = inv_xform_pt(m, pt);
  V3<R> mins_pt = mins - ls_pt;
  V3<R> pt_maxs = ls_pt - maxs;
  V3<R> zero(0, 0, 0);
  V3<R> shift = vmax(mins_pt, zero) - vmax(pt_maxs, zero);
  bool inside = (shift.x == R(0) && shift.y == R(0) && shift.z == R(0));
  V3<R> proj_ls;
  if (!inside) {
    proj_ls = ls_pt + shift;
  } else {
    R best = -Eps<R>::maxv();
    bool is_mins = false;
    int best_id = 0;
    for (int i = 0; i < 3; ++i) {
      R mins_pt_i = mins_pt[i], pt_maxs_i = pt_maxs[i];
      if (mins_pt_i < pt_maxs_i) {
        if (pt_maxs[i] > best) { best_id = i; is_mins = false; best = pt_maxs_i; }
      } else if (mins_pt_i > best) { best_id = i; is_mins = true; best = mins_pt_i; }
    }
    shift = zero;
    if (is_mins) shift[best_id] = best; else shift[best_id] = -best;
    proj_ls = ls_pt + shift;
  }
  *inside_out = inside;
  V3<R> proj = xform_pt(m, proj_ls);
  // feature (point_aabb.rs:79-134)
  int nzero = 0, last_zero = 0, last_not_zero = 0;
  for (int i = 0; i < 3; ++i) { if (shift[i] == R(0)) { nzero++; last_zero = i; } else last_not_zero = i; }
  V3<R> center = (mins + maxs) * R(0.5);
  if (nzero == 3) {
    *fkind = 3; *fid = 0;
    for (int i = 0; i < 3; ++i) {
      if (proj_ls[i] > maxs[i] - Eps<R>::eps()) { *fkind = 0; *fid = i; break; }
      if (proj_ls[i] <= mins[i] + Eps<R>::eps()) { *fkind = 0; *fid = i + 3; break; }
    }
  } else if (nzero == 2) {
    *fkind = 0;
    *fid = (proj_ls[last_not_zero] < center[last_not_zero]) ? last_not_zero + 3 : last_not_zero;
  } else {
    int id = 0;
    for (int i = 0; i < 3; ++i) if (proj_ls[i] < center[i]) id |= 1 << i;
    if (nzero == 0) { *fkind = 2; *fid = id; } else { *fkind = 1; *fid = (id << 2) | last_zero; }
  }
  return proj;
}

// Cuboid::feature_normal (cuboid.rs:511-561), local frame (the reference does not rotate it).
template <class R> inline V3<R> cuboid_feature_normal(int fkind, int fid) {
  V3<R> d(0, 0, 0);
  if (fkind == 0) {
    if (fid < 3) d[fid] = R(1); else d[fid - 3] = R(-1);
    return d;
  } else if (fkind == 1) {
    int edge = fid & 3, face1 = (edge + 1) % 3, face2 = (edge + 2) % 3, signs = fid >> 2;
    d[face1] = (signs & (1 << face1)) ? R(-1) : R(1);
    d[face2] = (signs & (1 << face2)) ? R(-1) : R(1);
    return normalize(d);
  } else {
    for (int i = 0; i < 3; ++i) d[i] = (fid & (1 << i)) ? R(-1) : R(1);
    return normalize(d);
  }
}

// approx::relative_eq!(x, y) with the default epsilon and max_relative (EPSILON of the type).
template <class R> inline bool relative_eq_r(R x, R y) {
  if (x == y) return true;
  if (std::isinf(x) || std::isinf(y)) return false;
  R d = std::fabs(x - y);
  if (d <= Eps<R>::eps()) return true;
  R ax = std::fabs(x), ay = std::fabs(y);
  return d <= (ay > ax ? ay : ax) * Eps<R>::eps();
}

// Triangle::project_point_with_location with solid = true for a triangle placed by m
// (query/point/point_triangle.rs:61-240); *inside = relative_eq!(projection, pt) (3D, :6-19), true
// on the solid branch.  Returns the projection in the world frame.
template <class R> inline V3<R> project_triangle_at(const Iso<R>& m, V3<R> a, V3<R> b, V3<R> c, V3<R> pt, bool* inside) {
  V3<R> p = inv_xform_pt(m, pt);
  V3<R> ab = b - a, ac = c - a, ap = p - a;
  R ab_ap = dot(ab, ap), ac_ap = dot(ac, ap);
  V3<R> res;
  bool done = false;
  if (ab_ap <= R(0) && ac_ap <= R(0)) { res = xform_pt(m, a); done = true; }
  V3<R> bp = p - b;
  R ab_bp = dot(ab, bp), ac_bp = dot(ac, bp);
  if (!done && ab_bp >= R(0) && ac_bp <= ab_bp) { res = xform_pt(m, b); done = true; }
  V3<R> cp = p - c;
  R ab_cp = dot(ab, cp), ac_cp = dot(ac, cp);
  if (!done && ac_cp >= R(0) && ab_cp <= ac_cp) { res = xform_pt(m, c); done = true; }
  if (!done) {
    V3<R> bc = c - b;
    V3<R> n = cross(ab, ac);
    R vc = dot(n, cross(ab, ap));
    R vb = -dot(n, cross(ac, cp));
    R va = dot(n, cross(bc, bp));
    if (vc < R(0) && ab_ap >= R(0) && ab_bp <= R(0)) res = xform_pt(m, a + ab * (ab_ap / norm_sq(ab)));
    else if (vb < R(0) && ac_ap >= R(0) && ac_cp <= R(0)) res = xform_pt(m, a + ac * (ac_ap / norm_sq(ac)));
    else if (va < R(0) && ac_bp - ab_bp >= R(0) && ab_cp - ac_cp >= R(0))
      res = xform_pt(m, b + bc * (dot(bc, bp) / norm_sq(bc)));
    else if (va + vb + vc != R(0)) {
      R denom = R(1) / (va + vb + vc);
      R v = vb * denom, w = vc * denom;
      res = xform_pt(m, a + ab * v + ac * w);
    } else {
      *inside = true;
      return pt;
    }
  }
  *inside = relative_eq_r(res.x, pt.x) && relative_eq_r(res.y, pt.y) && relative_eq_r(res.z, pt.z);
  return res;
}

// contact_ball_convex_polyhedron (contact_ball_convex_polyhedron.rs:11-58) for Cuboid and ConvexHull.
// For ConvexHull, the reference's zero-distance branch needs the hull's face/edge adjacency
// (convex.rs:388-415); this restatement returns None there (ball centre within eps of the surface).
template <class R> inline bool contact_ball_convex(V3<R> c1, R r, const Iso<R>& m2, const Shape<R>& g2, R prediction,
                                                   Contact<R>* out, int* fkind_out = nullptr, int* fid_out = nullptr,
                                                   bool* inside_out = nullptr) {
  bool inside;
  V3<R> world2;
  int fkind = 3, fid = 0;
  V3<R> tri_n(0, 1, 0);  // Triangle::feature_normal: its normal, or +y when degenerate (triangle.rs:367-374)
  if (g2.type == CUBOID) world2 = project_point_cuboid(m2, V3<R>(g2.a, g2.b, g2.c), c1, &inside, &fkind, &fid);
  else if (g2.type == SEGMENT) {
    // Segment::project_point_with_feature (query/point/point_segment.rs:13-91): kind 2 vertex
    // (id 0 or 1) or 1 edge (id 0), as project_point_cuboid's kinds
    V3<R> a(R(0), -g2.a, R(0)), b(R(0), g2.a, R(0));
    V3<R> ls = inv_xform_pt(m2, c1);
    V3<R> ab = b - a, ap = ls - a;
    R ab_ap = dot(ab, ap), sqnab = norm_sq(ab);
    if (ab_ap <= R(0)) { fkind = 2; fid = 0; world2 = xform_pt(m2, a); }
    else if (ab_ap >= sqnab) { fkind = 2; fid = 1; world2 = xform_pt(m2, b); }
    else { R u = ab_ap / sqnab; fkind = 1; fid = 0; world2 = xform_pt(m2, a + ab * u); }
    inside = relative_eq_r(world2.x, c1.x) && relative_eq_r(world2.y, c1.y) && relative_eq_r(world2.z, c1.z);
  } else if (g2.type == TRIANGLE) {
    world2 = project_triangle_at(m2, g2.pts[0], g2.pts[1], g2.pts[2], c1, &inside);
    V3<R> nn;
    if (try_normalize(cross(g2.pts[1] - g2.pts[0], g2.pts[2] - g2.pts[0]), Eps<R>::eps(), &nn)) tri_n = nn;
  } else world2 = project_point_support_map(m2, g2, c1, &inside);
  V3<R> dpt = world2 - c1;
  V3<R> dir; R dist;
  R depth; V3<R> normal;
  if (try_normalize(dpt, Eps<R>::eps(), &dir, &dist)) {
    if (inside) { depth = dist + r; normal = -dir; }
    else { depth = -dist + r; normal = dir; }
  } else if (g2.type == TRIANGLE) {  // every triangle location names a feature
    depth = r;
    normal = -tri_n;
  } else if (g2.type == SEGMENT) {  // Segment::feature_normal (shape/segment.rs:238-273), local frame
    V3<R> d;
    if (!try_normalize(V3<R>(R(0), g2.a, R(0)) - V3<R>(R(0), -g2.a, R(0)), Eps<R>::eps(), &d)) normal = V3<R>(0, 1, 0);
    else if (fkind == 2) normal = fid == 0 ? d : -d;
    else {
      int im = 0;  // Vector::iamin: first index of the smallest |component|
      R best = std::fabs(d.x);
      if (std::fabs(d.y) < best) { best = std::fabs(d.y); im = 1; }
      if (std::fabs(d.z) < best) im = 2;
      V3<R> nn(0, 0, 0);
      nn[im] = R(1);
      normal = normalize(nn - d * d[im]);
    }
    depth = r;
    normal = -normal;
  } else {
    if (g2.type != CUBOID || fkind == 3) return false;
    depth = r;
    normal = -cuboid_feature_normal<R>(fkind, fid);
  }
  if (fkind_out) *fkind_out = fkind;
  if (fid_out) *fid_out = fid;
  if (inside_out) *inside_out = inside;
  if (depth >= -prediction) {
    out->world1 = c1 + normal * r;
    out->world2 = world2;
    out->normal = normal;
    out->depth = depth;
    return true;
  }
  return false;
}

// query::contact (contact_shape_shape.rs:10-48) for {Ball, Cuboid, Capsule, ConvexHull}.
template <class R> inline bool contact(const Iso<R>& m1, const Shape<R>& g1, const Iso<R>& m2, const Shape<R>& g2,
                                       R prediction, Contact<R>* out, int* gjk_iters = nullptr, int* epa_iters = nullptr) {
  if (gjk_iters) *gjk_iters = 0;
  if (epa_iters) *epa_iters = 0;
  bool b1 = g1.type == BALL, b2 = g2.type == BALL;
  bool poly1 = g1.type == CUBOID || g1.type == HULL || g1.type == TRIANGLE;
  bool poly2 = g2.type == CUBOID || g2.type == HULL || g2.type == TRIANGLE;
  if (b1 && b2) return contact_ball_ball(m1.t, g1.a, m2.t, g2.a, prediction, out);
  if (b1 && poly2) return contact_ball_convex(m1.t, g1.a, m2, g2, prediction, out);
  if (poly1 && b2) {
    bool ok = contact_ball_convex(m2.t, g2.a, m1, g1, prediction, out);
    if (ok) { std::swap(out->world1, out->world2); out->normal = -out->normal; }  // Contact::flip
    return ok;
  }
  return contact_support_map_support_map(m1, g1, m2, g2, prediction, (const V3<R>*)nullptr, out, gjk_iters, epa_iters);
}

// query::distance (distance.rs:9-33).  init_dir / max_iter: the init_dir argument of
// distance_support_map_support_map_with_params with GJK capped at max_iter (< 10000) and plain
// GJK from the default direction on the cap; gjk_iters counts both runs.
template <class R> inline R distance(const Iso<R>& m1, const Shape<R>& g1, const Iso<R>& m2, const Shape<R>& g2,
                                     const V3<R>* init_dir = nullptr, int* gjk_iters = nullptr, int max_iter = 10000) {
  if (gjk_iters) *gjk_iters = 0;
  if (g1.type == BALL && g2.type == BALL) {
    // distance_ball_ball.rs:7-24
    V3<R> delta = m2.t - m1.t;
    R d2 = norm_sq(delta);
    R s = g1.a + g2.a;
    if (d2 <= s * s) return R(0);
    return std::sqrt(d2) - s;
  }
  // distance_support_map_support_map_with_params (distance_support_map_support_map.rs:25-53)
  int total = 0;
  for (int pass = 0; pass < 2; ++pass) {
    bool warm = init_dir && pass == 0;
    V3<R> d0 = warm ? *init_dir : (m1.t - m2.t);
    V3<R> dir;
    VoronoiSimplex<R> simplex;
    if (try_normalize(d0, Eps<R>::eps(), &dir)) simplex.reset(cso_from_shapes(m1, g1, m2, g2, dir));
    else simplex.reset(cso_from_shapes(m1, g1, m2, g2, V3<R>(1, 0, 0)));
    GJKResult<R> r = gjk_closest_points(m1, g1, m2, g2, Eps<R>::maxv(), true, simplex, warm ? max_iter : 10000);
    total += r.iters;
    if (gjk_iters) *gjk_iters = total;
    if (r.status == GJK_CAP) continue;
    if (r.status == GJK_CLOSEST) return norm(r.p1 - r.p2);
    return R(0);
  }
  return R(0);
}

// ---------------------------------------------------------------- AABBs
template <class R> struct AABB { V3<R> mins, maxs; };

template <class R> inline bool aabb_intersects(const AABB<R>& a, const AABB<R>& b) {  // aabb.rs:156-158
  return a.mins.x <= b.maxs.x && a.mins.y <= b.maxs.y && a.mins.z <= b.maxs.z &&
         a.maxs.x >= b.mins.x && a.maxs.y >= b.mins.y && a.maxs.z >= b.mins.z;
}
template <class R> inline bool aabb_contains(const AABB<R>& a, const AABB<R>& b) {  // aabb.rs:161-163
  return a.mins.x <= b.mins.x && a.mins.y <= b.mins.y && a.mins.z <= b.mins.z &&
         a.maxs.x >= b.maxs.x && a.maxs.y >= b.maxs.y && a.maxs.z >= b.maxs.z;
}
template <class R> inline AABB<R> aabb_merged(const AABB<R>& a, const AABB<R>& b) { return AABB<R>{vmin(a.mins, b.mins), vmax(a.maxs, b.maxs)}; }
template <class R> inline AABB<R> aabb_loosened(const AABB<R>& a, R m) {
  return AABB<R>{a.mins + V3<R>(-m, -m, -m), a.maxs + V3<R>(m, m, m)};
}
template <class R> inline V3<R> aabb_center(const AABB<R>& a) { return (a.mins + a.maxs) * R(0.5); }

// Shape::aabb for the four shapes (aabb_ball.rs:8-13, aabb_cuboid.rs:7-14, aabb_support_map.rs:35-45
// -> aabb_utils.rs:9-31, aabb_convex.rs:7-11 -> aabb_utils.rs:59-79).
template <class R> inline AABB<R> shape_aabb(const Shape<R>& g, const Iso<R>& m) {
  AABB<R> bb;
  switch (g.type) {
    case BALL:
      bb.mins = m.t + V3<R>(-g.a, -g.a, -g.a);
      bb.maxs = m.t + V3<R>(g.a, g.a, g.a);
      return bb;
    case CUBOID: {
      R M[3][3];
      quat_to_mat(m.q, M);
      V3<R> h(g.a, g.b, g.c), ws;
      for (int r = 0; r < 3; ++r) ws[r] = (std::fabs(M[r][0]) * h.x + std::fabs(M[r][1]) * h.y) + std::fabs(M[r][2]) * h.z;
      bb.mins = m.t - ws;
      bb.maxs = m.t + ws;
      return bb;
    }
    case CAPSULE: {
      V3<R> mn, mx;
      for (int d = 0; d < 3; ++d) {
        V3<R> basis(0, 0, 0);
        basis[d] = R(1);
        mx[d] = support_point(g, m, basis)[d];
        basis[d] = R(-1);
        mn[d] = support_point(g, m, basis)[d];
      }
      bb.mins = mn; bb.maxs = mx;
      return bb;
    }
    case HULL: {
      V3<R> w0 = xform_pt(m, g.pts[0]);
      V3<R> mn = w0, mx = w0;
      for (uint32_t i = 1; i < g.npts; ++i) { V3<R> w = xform_pt(m, g.pts[i]); mn = vmin(mn, w); mx = vmax(mx, w); }
      bb.mins = mn; bb.maxs = mx;
      return bb;
    }
  }
  bb.mins = bb.maxs = m.t;
  return bb;
}

// ---------------------------------------------------------------- ray casting
template <class R> struct Ray { V3<R> origin, dir; };

// AABB::toi_with_ray (ray_aabb.rs:13-50), identity isometry.
template <class R> inline bool ray_aabb_toi(const AABB<R>& bb, const Ray<R>& ray, R max_toi, bool solid, R* toi) {
  R tmin = R(0), tmax = max_toi;
  for (int i = 0; i < 3; ++i) {
    if (ray.dir[i] == R(0)) {
      if (ray.origin[i] < bb.mins[i] || ray.origin[i] > bb.maxs[i]) return false;
    } else {
      R denom = R(1) / ray.dir[i];
      R tn = (bb.mins[i] - ray.origin[i]) * denom;
      R tf = (bb.maxs[i] - ray.origin[i]) * denom;
      if (tn > tf) std::swap(tn, tf);
      tmin = std::fmax(tmin, tn);
      tmax = std::fmin(tmax, tf);
      if (tmin > tmax) return false;
    }
  }
  *toi = (tmin == R(0) && !solid) ? tmax : tmin;
  return true;
}

// ray_intersection_with_triangle (ray_triangle.rs:32-114); fid 0 = front, 1 = back.
template <class R> inline bool ray_triangle(V3<R> a, V3<R> b, V3<R> c, const Ray<R>& ray, R* toi_out, V3<R>* n_out, int* fid_out) {
  V3<R> ab = b - a, ac = c - a;
  V3<R> n = cross(ab, ac);
  R d = dot(n, ray.dir);
  if (d == R(0)) return false;
  V3<R> ap = ray.origin - a;
  R t = dot(ap, n);
  if ((t < R(0) && d < R(0)) || (t > R(0) && d > R(0))) return false;
  int fid = d < R(0) ? 0 : 1;
  d = std::fabs(d);
  V3<R> e = -cross(ray.dir, ap);
  R v, w, toi;
  V3<R> normal;
  if (t < R(0)) {
    v = -dot(ac, e);
    if (v < R(0) || v > d) return false;
    w = dot(ab, e);
    if (w < R(0) || v + w > d) return false;
    R invd = R(1) / d;
    toi = -t * invd;
    normal = -normalize(n);
  } else {
    v = dot(ac, e);
    if (v < R(0) || v > d) return false;
    w = -dot(ab, e);
    if (w < R(0) || v + w > d) return false;
    R invd = R(1) / d;
    toi = t * invd;
    normal = normalize(n);
  }
  *toi_out = toi; *n_out = normal; *fid_out = fid;
  return true;
}

// ---------------------------------------------------------------- BVT (partitioning/bvt.rs)
template <class R> struct BVT {
  struct Node { bool leaf; uint32_t idx; };
  struct Internal { AABB<R> bv; Node left, right; };
  struct Leaf { AABB<R> bv; uint32_t data; };
  std::vector<Internal> internals;
  std::vector<Leaf> leaves;
  Node root{true, 0};

  // utils/median.rs:5-17
  static R median(std::vector<R>& v) {
    std::sort(v.begin(), v.end());
    size_t n = v.size();
    if (n % 2 == 0) return (v[n / 2 - 1] + v[n / 2]) / R(2);
    return v[n / 2];
  }
  // median_partitioning_with_centers (bvt.rs:290-350) + _from_partitioning (:352-404)
  Node build(size_t depth, std::vector<std::pair<uint32_t, AABB<R>>>&& elts) {
    if (elts.size() == 1) {
      leaves.push_back(Leaf{elts[0].second, elts[0].first});
      return Node{true, (uint32_t)leaves.size() - 1};
    }
    int axis = (int)(depth % 3);
    std::vector<R> med;
    med.reserve(elts.size());
    for (auto& l : elts) med.push_back(aabb_center(l.second)[axis]);
    R m = median(med);
    std::vector<std::pair<uint32_t, AABB<R>>> left, right;
    AABB<R> bb = elts[0].second;
    bool insert_left = false;
    for (auto& l : elts) {
      bb = aabb_merged(bb, l.second);
      R pos = aabb_center(l.second)[axis];
      if (pos < m || (pos == m && insert_left)) { left.push_back(l); insert_left = false; }
      else { right.push_back(l); insert_left = true; }
    }
    if (left.empty()) { left.push_back(right.back()); right.pop_back(); }
    else if (right.empty()) { right.push_back(left.back()); left.pop_back(); }
    std::vector<std::pair<uint32_t, AABB<R>>>().swap(elts);
    Node l = build(depth + 1, std::move(left));
    Node r = build(depth + 1, std::move(right));
    internals.push_back(Internal{bb, l, r});
    return Node{false, (uint32_t)internals.size() - 1};
  }
  void build_balanced(std::vector<std::pair<uint32_t, AABB<R>>>&& elts) {
    internals.clear(); leaves.clear();
    if (elts.empty()) return;
    root = build(0, std::move(elts));
  }
  const AABB<R>& bv(Node n) const { return n.leaf ? leaves[n.idx].bv : internals[n.idx].bv; }
};

// TriMesh: points + indices, BVT over local triangle AABBs (trimesh.rs:100-197).
template <class R> struct TriMesh {
  std::vector<V3<R>> points;
  std::vector<uint32_t> idx;  // 3 per face
  BVT<R> bvt;
  void build() {
    size_t nf = idx.size() / 3;
    std::vector<std::pair<uint32_t, AABB<R>>> leaves;
    leaves.reserve(nf);
    for (size_t i = 0; i < nf; ++i) {
      V3<R> a = points[idx[3 * i]], b = points[idx[3 * i + 1]], c = points[idx[3 * i + 2]];
      AABB<R> bb;  // aabb_triangle.rs:27-41
      for (int d = 0; d < 3; ++d) {
        bb.mins[d] = std::fmin(std::fmin(a[d], b[d]), c[d]);
        bb.maxs[d] = std::fmax(std::fmax(a[d], b[d]), c[d]);
      }
      leaves.push_back({(uint32_t)i, bb});
    }
    bvt.build_balanced(std::move(leaves));
  }
  struct Hit { R toi; V3<R> normal; uint32_t feature; };
  struct Weighted { typename BVT<R>::Node node; R cost; };
  struct WLE { bool operator()(const Weighted& a, const Weighted& b) const { return a.cost <= b.cost; } };

  // TriMesh::toi_and_normal_with_ray (ray_trimesh.rs:23-50) -> BVH::best_first_search (bvh.rs:101-160)
  // with TriMeshRayToiAndNormalVisitor (ray_trimesh.rs:150-191). Mesh at `m`.
  bool cast(const Iso<R>& m, const Ray<R>& wray, R max_toi, Hit* hit, uint32_t* nodes_visited = nullptr) const {
    if (bvt.leaves.empty()) return false;
    Ray<R> ray{inv_xform_pt(m, wray.origin), inv_xform_vec(m, wray.dir)};
    uint32_t nf = (uint32_t)(idx.size() / 3);
    uint32_t visits = 0;
    auto visit = [&](R best, typename BVT<R>::Node n, R* cost, bool* has_res, Hit* res) -> bool {
      visits++;
      R toi;
      if (!ray_aabb_toi(bvt.bv(n), ray, max_toi, true, &toi)) return false;  // Stop
      *cost = toi;
      *has_res = false;
      if (n.leaf && toi < best) {
        uint32_t f = bvt.leaves[n.idx].data;
        V3<R> a = points[idx[3 * f]], b = points[idx[3 * f + 1]], c = points[idx[3 * f + 2]];
        R t; V3<R> nn; int fid;
        if (ray_triangle(a, b, c, ray, &t, &nn, &fid) && t <= max_toi) {
          *cost = t;
          *has_res = true;
          res->toi = t;
          res->normal = nn;
          res->feature = fid == 1 ? f + nf : f;
        }
      }
      return true;
    };
    RustHeap<Weighted, WLE> queue;
    R best_cost = Eps<R>::maxv();
    bool found = false;
    Hit best_res{};
    R cost; bool has; Hit r;
    if (!visit(best_cost, bvt.root, &cost, &has, &r)) { if (nodes_visited) *nodes_visited = visits; return false; }
    if (has) { best_cost = cost; best_res = r; found = true; }
    queue.push(Weighted{bvt.root, -cost});
    Weighted e;
    while (queue.pop(&e)) {
      if (-e.cost >= best_cost) break;
      if (e.node.leaf) continue;
      const auto& in = bvt.internals[e.node.idx];
      typename BVT<R>::Node ch[2] = {in.left, in.right};
      for (int i = 0; i < 2; ++i) {
        if (visit(best_cost, ch[i], &cost, &has, &r)) {
          if (cost < best_cost) {
            if (has) { best_cost = cost; best_res = r; found = true; }
            queue.push(Weighted{ch[i], -cost});
          }
        }
      }
    }
    if (nodes_visited) *nodes_visited = visits;
    if (!found) return false;
    best_res.normal = xform_vec(m, best_res.normal);
    *hit = best_res;
    return true;
  }
};

// ---------------------------------------------------------------- slab allocator (slab crate: LIFO free list)
template <class T> struct Slab {
  struct Entry { bool occupied; T val; size_t next; };
  std::vector<Entry> entries;
  size_t next = 0, len = 0;
  size_t insert(const T& v) {
    size_t key = next;
    if (key == entries.size()) { entries.push_back(Entry{true, v, 0}); next = key + 1; }
    else { next = entries[key].next; entries[key].occupied = true; entries[key].val = v; }
    len++;
    return key;
  }
  T remove(size_t key) {
    T v = entries[key].val;
    entries[key].occupied = false;
    entries[key].next = next;
    next = key;
    len--;
    return v;
  }
  void clear() { entries.clear(); next = 0; len = 0; }
  bool contains(size_t k) const { return k < entries.size() && entries[k].occupied; }
  T& operator[](size_t k) { return entries[k].val; }
  const T& operator[](size_t k) const { return entries[k].val; }
};

// ---------------------------------------------------------------- DBVT (partitioning/dbvt.rs)
template <class R> struct DBVT {
  enum PKind { ROOT, LEFT_OF, RIGHT_OF };
  struct PId { PKind kind; size_t p; };
  struct NId { bool leaf; size_t id; };
  struct Leaf { AABB<R> bv; V3<R> center; uint32_t data; PId parent; };
  struct Internal { AABB<R> bv; V3<R> center; NId left, right; PId parent; };
  NId root{true, 0};
  Slab<Leaf> leaves;
  Slab<Internal> internals;

  V3<R> center_of(NId n) const { return n.leaf ? leaves[n.id].center : internals[n.id].center; }
  // dbvt.rs:158-255
  size_t insert(AABB<R> bv, uint32_t data) {
    Leaf leaf{bv, aabb_center(bv), data, PId{ROOT, 0}};
    if (leaves.len == 0) {
      size_t nid = leaves.insert(leaf);
      leaves[nid].parent = PId{ROOT, 0};
      root = NId{true, nid};
      return nid;
    }
    if (!root.leaf) {
      NId curr = root;
      for (;;) {
        if (!curr.leaf) {
          Internal& node = internals[curr.id];
          node.bv = aabb_merged(node.bv, leaf.bv);
          NId l = node.left, r = node.right;
          V3<R> d1v = center_of(l) - leaf.center;
          V3<R> d2v = center_of(r) - leaf.center;
          R d1 = norm_sq(d1v), d2 = norm_sq(d2v);
          curr = d1 < d2 ? l : r;
        } else {
          size_t id = curr.id;
          AABB<R> pbv = aabb_merged(leaves[id].bv, leaf.bv);
          PId gp = leaves[id].parent;
          size_t nid = leaves.insert(leaf);
          Internal par{pbv, aabb_center(pbv), curr, NId{true, nid}, gp};
          size_t pid = internals.insert(par);
          leaves[id].parent = PId{LEFT_OF, pid};
          leaves[nid].parent = PId{RIGHT_OF, pid};
          if (gp.kind == LEFT_OF) internals[gp.p].left = NId{false, pid};
          else internals[gp.p].right = NId{false, pid};
          return nid;
        }
      }
    } else {
      size_t id = root.id;
      size_t nid = leaves.insert(leaf);
      AABB<R> rbv = aabb_merged(leaves[id].bv, leaves[nid].bv);
      Internal rt{rbv, aabb_center(rbv), NId{true, id}, NId{true, nid}, PId{ROOT, 0}};
      size_t rid = internals.insert(rt);
      leaves[id].parent = PId{LEFT_OF, rid};
      leaves[nid].parent = PId{RIGHT_OF, rid};
      root = NId{false, rid};
      return nid;
    }
  }
  // dbvt.rs:260-328
  Leaf remove(size_t leaf_id) {
    Leaf leaf = leaves.remove(leaf_id);
    if (leaf.parent.kind != ROOT) {
      size_t p = leaf.parent.p;
      NId other = leaf.parent.kind == RIGHT_OF ? internals[p].left : internals[p].right;
      PId pp = internals[p].parent;
      if (pp.kind == ROOT) {
        if (other.leaf) leaves[other.id].parent = PId{ROOT, 0}; else internals[other.id].parent = PId{ROOT, 0};
        root = other;
      } else {
        if (other.leaf) leaves[other.id].parent = PId{pp.kind, pp.p}; else internals[other.id].parent = PId{pp.kind, pp.p};
        if (pp.kind == RIGHT_OF) internals[pp.p].right = other; else internals[pp.p].left = other;
      }
      internals.remove(p);
    } else {
      leaves.clear();
      internals.clear();
    }
    return leaf;
  }
  // BVH::visit (bvh.rs:24-45) + BoundingVolumeInterferencesCollector (bounding_volume_interferences_collector.rs:41-51)
  void collect(const AABB<R>& q, std::vector<uint32_t>& out) const {
    if (leaves.len == 0) return;
    std::vector<NId> stack;
    stack.push_back(root);
    while (!stack.empty()) {
      NId n = stack.back();
      stack.pop_back();
      const AABB<R>& bv = n.leaf ? leaves[n.id].bv : internals[n.id].bv;
      if (aabb_intersects(bv, q)) {
        if (n.leaf) out.push_back(leaves[n.id].data);
        else { stack.push_back(internals[n.id].left); stack.push_back(internals[n.id].right); }
      }
    }
  }
};

// ---------------------------------------------------------------- DBVTBroadPhase (pipeline/broad_phase/dbvt_broad_phase.rs)
template <class R> struct DBVTBroadPhase {
  enum Status { DETACHED_NONE, DETACHED_SOME, ON_DYNAMIC, ON_STATIC, DELETED };
  struct Proxy { uint32_t data; Status status; size_t leaf; size_t energy; bool updated; };
  static const size_t DEACTIVATION_THRESHOLD = 100;
  Slab<Proxy> proxies;
  DBVT<R> tree, stree;
  // pair set keyed by (min, max) handle; value = up_to_date flag
  std::vector<std::pair<uint64_t, bool>> pairs;  // kept unordered like a HashMap; order is not part of the contract
  R margin;
  struct PendingLeaf { AABB<R> bv; uint32_t handle; bool valid; };
  std::vector<PendingLeaf> leaves_to_update;
  std::vector<std::pair<uint32_t, AABB<R>>> proxies_to_update;  // VecDeque (push_back order)
  std::function<bool(uint32_t, uint32_t)> allowed;
  std::vector<uint64_t> started, stopped;

  explicit DBVTBroadPhase(R m) : margin(m) {}
  static uint64_t key(uint32_t a, uint32_t b) { return a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a; }
  size_t find_pair(uint64_t k) const {
    for (size_t i = 0; i < pairs.size(); ++i) if (pairs[i].first == k) return i;
    return (size_t)-1;
  }

  uint32_t create_proxy(const AABB<R>& bv) {  // :275-280 (stores the box as given)
    Proxy p{0, DETACHED_NONE, 0, 0, true};
    uint32_t h = (uint32_t)proxies.insert(p);
    proxies[h].data = h;
    proxies_to_update.push_back({h, bv});
    return h;
  }
  void deferred_set_bounding_volume(uint32_t h, const AABB<R>& bv) {  // :325-347
    Proxy& p = proxies[h];
    bool needs;
    if (p.status == ON_STATIC) needs = !aabb_contains(stree.leaves[p.leaf].bv, bv);
    else if (p.status == ON_DYNAMIC) needs = !aabb_contains(tree.leaves[p.leaf].bv, bv);
    else needs = true;
    if (needs) proxies_to_update.push_back({h, aabb_loosened(bv, margin)});
  }
  const AABB<R>* stored_bv(uint32_t h) const {
    const Proxy& p = proxies[h];
    if (p.status == ON_STATIC) return &stree.leaves[p.leaf].bv;
    if (p.status == ON_DYNAMIC) return &tree.leaves[p.leaf].bv;
    return nullptr;
  }
  void remove(const std::vector<uint32_t>& hs) {  // :282-323
    for (uint32_t h : hs) {
      Proxy& p = proxies[h];
      if (p.status == ON_STATIC) stree.remove(p.leaf);
      else if (p.status == ON_DYNAMIC) tree.remove(p.leaf);
      p.status = DELETED;
    }
    std::vector<std::pair<uint64_t, bool>> kept;
    for (auto& pr : pairs) {
      uint32_t a = (uint32_t)(pr.first >> 32), b = (uint32_t)pr.first;
      if (proxies[a].status == DELETED || proxies[b].status == DELETED) stopped.push_back(pr.first);
      else kept.push_back(pr);
    }
    pairs.swap(kept);
    for (uint32_t h : hs) proxies.remove(h);
  }
  void update() {  // :174-259
    started.clear();
    for (auto& pu : proxies_to_update) {
      uint32_t h = pu.first;
      if (!proxies.contains(h)) continue;
      Proxy& p = proxies[h];
      bool set_status = true;
      if (p.status == ON_STATIC) { stree.remove(p.leaf); leaves_to_update.push_back({pu.second, h, true}); }
      else if (p.status == ON_DYNAMIC) { tree.remove(p.leaf); leaves_to_update.push_back({pu.second, h, true}); }
      else if (p.status == DETACHED_NONE) { leaves_to_update.push_back({pu.second, h, true}); }
      else if (p.status == DETACHED_SOME) { leaves_to_update[p.leaf] = {pu.second, h, true}; set_status = false; }
      p.updated = true;
      if (set_status) { p.status = DETACHED_SOME; p.leaf = leaves_to_update.size() - 1; }
    }
    proxies_to_update.clear();
    bool some = !leaves_to_update.empty();
    std::vector<uint32_t> collector;
    for (auto& lf : leaves_to_update) {
      collector.clear();
      tree.collect(lf.bv, collector);
      stree.collect(lf.bv, collector);
      for (uint32_t h2 : collector) {
        if (allowed && !allowed(lf.handle, h2)) continue;
        uint64_t k = key(lf.handle, h2);
        size_t at = find_pair(k);
        if (at != (size_t)-1) pairs[at].second = true;
        else { started.push_back(k); pairs.push_back({k, true}); }
      }
      size_t leaf = tree.insert(lf.bv, lf.handle);
      proxies[lf.handle].status = ON_DYNAMIC;
      proxies[lf.handle].leaf = leaf;
      proxies[lf.handle].energy = DEACTIVATION_THRESHOLD;
    }
    leaves_to_update.clear();
    if (some) {  // purge_some_contact_pairs (:101-147), purge_all = false
      std::vector<std::pair<uint64_t, bool>> kept;
      for (auto& pr : pairs) {
        bool retain = true;
        if (!pr.second) {
          uint32_t a = (uint32_t)(pr.first >> 32), b = (uint32_t)pr.first;
          if (proxies[a].updated || proxies[b].updated) {
            if (!allowed || allowed(a, b)) {
              if (!aabb_intersects(*stored_bv(a), *stored_bv(b))) { stopped.push_back(pr.first); retain = false; }
            } else { stopped.push_back(pr.first); retain = false; }
          }
        }
        pr.second = false;
        if (retain) kept.push_back(pr);
      }
      pairs.swap(kept);
    }
    // update_activation_states (:149-165)
    for (size_t h = 0; h < proxies.entries.size(); ++h) {
      if (!proxies.entries[h].occupied) continue;
      Proxy& p = proxies.entries[h].val;
      if (p.status == ON_DYNAMIC) {
        if (p.energy == 1) {
          typename DBVT<R>::Leaf old = tree.remove(p.leaf);
          size_t nl = stree.insert(old.bv, old.data);
          p.status = ON_STATIC; p.leaf = nl;
        } else p.energy -= 1;
      }
    }
  }
};


// ================================================================ time of impact and ray casts on shapes
// approx::relative_eq!(x, 0) with the default epsilon and max_relative (both EPSILON of the type):
// |x - 0| <= EPSILON, and the relative test |x| <= |x| * EPSILON never holds for finite x != 0.
template <class R> inline bool relative_eq_zero(R x) { return std::fabs(x) <= Eps<R>::eps(); }

// SupportMap::support_point_toward (shape/support_map.rs:31-34): `u` is unit and used as given;
// Ball overrides it (ball.rs:36-39), Capsule's local_support_point_toward does not normalise
// (capsule.rs:79-83), Cuboid and ConvexHull fall back to support_point.
template <class R> inline V3<R> support_point_toward(const Shape<R>& g, const Iso<R>& m, V3<R> u) {
  switch (g.type) {
    case BALL: return m.t + u * g.a;
    case CAPSULE: {
      V3<R> ld = inv_xform_vec(m, u);
      V3<R> res(R(0), std::copysign(g.a, ld.y), R(0));
      return xform_pt(m, res + ld * g.b);
    }
    default: return support_point(g, m, u);
  }
}

// A support map, dilated by `radius` when `dilated`: DilatedShape (special_support_maps.rs:32-59),
// whose support_point normalises the direction and adds the ball's support to the shape's.
template <class R> struct SMap {
  const Shape<R>* g;
  R radius;
  bool dilated;
};
template <class R> inline SMap<R> smap(const Shape<R>& g) { return SMap<R>{&g, R(0), false}; }
template <class R> inline V3<R> smap_support(const SMap<R>& s, const Iso<R>& m, V3<R> dir) {
  if (!s.dilated) return support_point(*s.g, m, dir);
  V3<R> u = normalize(dir);
  return support_point_toward(*s.g, m, u) + u * s.radius;
}
template <class R> inline CSOPoint<R> cso_from_smaps(const Iso<R>& m1, const SMap<R>& g1, const Iso<R>& m2,
                                                     const SMap<R>& g2, V3<R> dir) {
  CSOPoint<R> c;
  c.orig1 = smap_support(g1, m1, dir);
  c.orig2 = smap_support(g2, m2, -dir);
  c.point = c.orig1 - c.orig2;
  return c;
}
template <class R> inline CSOPoint<R> cso_translated(const CSOPoint<R>& c, V3<R> d) {  // cso_point.rs:88-90
  CSOPoint<R> r = c;
  r.point = c.point + d;
  return r;
}

// gjk::minkowski_ray_cast (gjk.rs:245-365) on m1*g1 - m2*g2; false is None.  *iters counts the
// support evaluations.  The improved_fixed_point_support feature is off (the crate's default).
template <class R> inline bool minkowski_ray_cast(const Iso<R>& m1, const SMap<R>& g1, const Iso<R>& m2,
                                                  const SMap<R>& g2, const Ray<R>& ray, R max_toi,
                                                  VoronoiSimplex<R>& simplex, R* toi, V3<R>* normal,
                                                  int* iters = nullptr) {
  const R eps_tol = gjk_eps_tol<R>();
  const R eps_rel = std::sqrt(eps_tol);
  const R maxv = Eps<R>::maxv();
  int it = 0;
  if (iters) *iters = 0;
  R ray_length = norm(ray.dir);
  if (relative_eq_zero(ray_length)) return false;
  R ltoi = R(0);
  V3<R> co = ray.origin, cd = ray.dir / ray_length;
  V3<R> dir = -cd;
  V3<R> ldir = dir;
  CSOPoint<R> sp = cso_from_smaps(m1, g1, m2, g2, dir);
  it++;
  simplex.reset(cso_translated(sp, -co));
  V3<R> proj = simplex.project_origin_and_reduce();
  R max_bound = maxv;
  int niter = 0;
  bool last_chance = false;
  auto hit = [&]() {
    *toi = ltoi / ray_length;
    *normal = ldir;
    if (iters) *iters = it;
    return true;
  };
  auto miss = [&]() {
    if (iters) *iters = it;
    return false;
  };
  for (;;) {
    R old_max_bound = max_bound;
    {
      V3<R> nd;
      R dist;
      if (try_normalize(-proj, eps_tol, &nd, &dist)) { dir = nd; max_bound = dist; }
      else return hit();
    }
    if (max_bound >= old_max_bound) {
      last_chance = true;  // CSOPoint::single_point (cso_point.rs:42-44)
      V3<R> p = proj + co;
      sp.point = p; sp.orig1 = p; sp.orig2 = V3<R>(0, 0, 0);
    } else {
      sp = cso_from_smaps(m1, g1, m2, g2, dir);
      it++;
    }
    if (last_chance && ltoi > R(0)) return hit();
    // query::ray_toi_with_plane(support point, dir, curr_ray) (ray_plane.rs:8-37)
    R denom = dot(dir, cd);
    bool has_t = false;
    R t = R(0);
    if (!relative_eq_zero(denom)) {
      t = dot(dir, sp.point - co) / denom;
      has_t = t >= R(0);
    }
    if (has_t) {
      if (dot(dir, cd) < R(0) && t > R(0)) {
        ldir = dir;
        ltoi += t;
        if (ltoi / ray_length > max_toi) return miss();
        V3<R> shift = cd * t;
        co = co + shift;
        max_bound = maxv;
        for (int i = 0; i < simplex.dim + 1; ++i) simplex.vertices[i].point = simplex.vertices[i].point + (-shift);
        last_chance = false;
      }
    } else if (dot(dir, cd) > eps_tol) {
      return miss();
    }
    if (last_chance) return miss();
    R min_bound = -dot(dir, sp.point - co);
    if (max_bound - min_bound <= eps_rel * max_bound) return miss();
    (void)simplex.add_point(cso_translated(sp, -co));
    proj = simplex.project_origin_and_reduce();
    if (simplex.dim == 3) {
      if (min_bound >= eps_tol) return miss();
      return hit();
    }
    niter += 1;
    if (niter == 10000) return miss();
  }
}

// TOI status (time_of_impact.rs:7-25).
enum TOIStatus { TOI_OUT_OF_ITERATIONS = 0, TOI_CONVERGED = 1, TOI_FAILED = 2, TOI_PENETRATING = 3 };
template <class R> struct TOIResult {
  R toi;
  V3<R> witness1, witness2, normal1, normal2;  // local frames of shape 1 and 2
  int status;
};

// ray_toi_with_ball (ray_ball.rs:83-124); *inside reports the origin inside the ball.
template <class R> inline bool ray_toi_with_ball(V3<R> center, R radius, const Ray<R>& ray, bool solid, bool* inside,
                                                 R* toi) {
  V3<R> dcenter = ray.origin - center;
  R a = norm_sq(ray.dir);
  R b = dot(dcenter, ray.dir);
  R c = norm_sq(dcenter) - radius * radius;
  *inside = false;
  if (a == R(0)) {
    if (c > R(0)) return false;
    *inside = true;
    *toi = R(0);
    return true;
  }
  if (c > R(0) && b > R(0)) return false;
  R delta = b * b - a * c;
  if (delta < R(0)) return false;
  R t = (-b - std::sqrt(delta)) / a;
  if (t <= R(0)) {
    *inside = true;
    *toi = solid ? R(0) : (-b + std::sqrt(delta)) / a;
    return true;
  }
  *toi = t;
  return true;
}

// time_of_impact_ball_ball (time_of_impact_ball_ball.rs:7-61) followed by the frame change of
// query::time_of_impact (time_of_impact.rs:75-98).
template <class R> inline bool toi_ball_ball(const Iso<R>& m1, V3<R> vel1, R r1, const Iso<R>& m2, V3<R> vel2, R r2,
                                             R max_toi, R target_distance, TOIResult<R>* out) {
  V3<R> vel = vel1 - vel2;
  R rsum = r1 + r2;
  R radius = rsum + target_distance;
  V3<R> center = m1.t + (-m2.t);
  Ray<R> ray{V3<R>(0, 0, 0), -vel};
  bool inside;
  R toi;
  if (!ray_toi_with_ball(center, radius, ray, true, &inside, &toi)) return false;
  if (toi > max_toi) return false;
  V3<R> dpt = (ray.origin + ray.dir * toi) - center;
  V3<R> n1, n2, w1, w2;
  if (radius == R(0)) {
    n1 = V3<R>(1, 0, 0); n2 = V3<R>(1, 0, 0); w1 = V3<R>(0, 0, 0); w2 = V3<R>(0, 0, 0);
  } else {
    n1 = dpt / radius;
    n2 = -n1;
    w1 = n1 * r1;
    w2 = n2 * r2;
  }
  out->toi = toi;
  out->status = (inside && norm_sq(center) < rsum * rsum) ? TOI_PENETRATING : TOI_CONVERGED;
  out->witness1 = qrot_inv(m1.q, w1);
  out->witness2 = qrot_inv(m2.q, w2);
  out->normal1 = qrot_inv(m1.q, n1);
  out->normal2 = qrot_inv(m2.q, n2);
  return true;
}

// time_of_impact_support_map_support_map (time_of_impact_support_map_support_map.rs:10-116)
// through gjk::directional_distance (gjk.rs:214-242).
template <class R> inline bool toi_support_map_support_map(const Iso<R>& m1, V3<R> vel1, const Shape<R>& g1,
                                                           const Iso<R>& m2, V3<R> vel2, const Shape<R>& g2,
                                                           R max_toi, R target_distance, TOIResult<R>* out,
                                                           int* iters = nullptr) {
  V3<R> dvel = vel2 - vel1;
  SMap<R> s1 = smap(g1), s2 = smap(g2);
  if (target_distance != R(0)) { s1.radius = target_distance; s1.dilated = true; }
  VoronoiSimplex<R> simplex;
  Ray<R> ray{V3<R>(0, 0, 0), dvel};
  R toi;
  V3<R> normal;
  if (!minkowski_ray_cast(m1, s1, m2, s2, ray, Eps<R>::maxv(), simplex, &toi, &normal, iters)) return false;
  V3<R> w1(0, 0, 0), w2(0, 0, 0);
  if (toi != R(0)) gjk_result(simplex, simplex.dim == 3, &w1, &w2);
  if (toi > max_toi) return false;
  out->toi = toi;
  if (target_distance == R(0)) {
    out->normal1 = inv_xform_vec(m1, normal);
    out->normal2 = inv_xform_vec(m2, -normal);
    out->witness1 = inv_xform_pt(m1, w1);
    out->witness2 = inv_xform_pt(m2, w2);
    out->status = toi == R(0) ? TOI_PENETRATING : TOI_CONVERGED;
    return true;
  }
  int status = TOI_CONVERGED;
  if (toi == R(0)) {
    // closest_points_support_map_support_map_with_params(.., target_distance, Some(normal))
    // (closest_points_support_map_support_map.rs:40-67)
    VoronoiSimplex<R> s2x;
    V3<R> d;
    if (try_normalize(normal, Eps<R>::eps(), &d)) s2x.reset(cso_from_shapes(m1, g1, m2, g2, d));
    else s2x.reset(cso_from_shapes(m1, g1, m2, g2, V3<R>(1, 0, 0)));
    GJKResult<R> r = gjk_closest_points(m1, g1, m2, g2, target_distance, true, s2x);
    if (r.status == GJK_CLOSEST) {
      V3<R> nn = normalize(r.p2 - r.p1);
      out->normal1 = inv_xform_vec(m1, nn);
      out->normal2 = inv_xform_vec(m2, -nn);
      out->witness1 = inv_xform_pt(m1, w1);
      out->witness2 = inv_xform_pt(m2, w2);
      out->status = TOI_CONVERGED;
      return true;
    } else if (r.status == GJK_INTERSECTION) {
      status = TOI_PENETRATING;
    }
  }
  V3<R> n1 = inv_xform_vec(m1, normal);
  out->normal1 = n1;
  out->normal2 = inv_xform_vec(m2, -normal);
  out->witness1 = inv_xform_pt(m1, w1) - n1 * target_distance;
  out->witness2 = inv_xform_pt(m2, w2);
  out->status = status;
  return true;
}

// query::time_of_impact (time_of_impact.rs:66-158) for {Ball, Cuboid, Capsule, ConvexHull}:
// Ball-Ball closed form, every other pair through the support maps.
template <class R> inline bool time_of_impact(const Iso<R>& m1, V3<R> vel1, const Shape<R>& g1, const Iso<R>& m2,
                                              V3<R> vel2, const Shape<R>& g2, R max_toi, R target_distance,
                                              TOIResult<R>* out, int* iters = nullptr) {
  if (iters) *iters = 0;
  if (g1.type == BALL && g2.type == BALL)
    return toi_ball_ball(m1, vel1, g1.a, m2, vel2, g2.a, max_toi, target_distance, out);
  return toi_support_map_support_map(m1, vel1, g1, m2, vel2, g2, max_toi, target_distance, out, iters);
}

// RayIntersection of a shape: toi, world normal, feature (FEAT_FACE0 + face id; FEAT_UNKNOWN for
// FeatureId::Unknown).
constexpr uint32_t FEAT_UNKNOWN = 0xFFFFFFFEu;
template <class R> struct RayHit {
  R toi;
  V3<R> normal;
  uint32_t feature;
};

// AABB ray cast with the entry/exit normals: clip_line (ray_aabb.rs:178-268) and ray_aabb
// (:270-290), local frame.  *side: +(i+1) / -(i+1) as in the reference.
template <class R> inline bool ray_aabb_normal(const AABB<R>& bb, const Ray<R>& ray, R max_toi, bool solid, R* toi,
                                               V3<R>* normal, int* side) {
  R tmax = Eps<R>::maxv();
  R tmin = -tmax;
  int near_side = 0, far_side = 0;
  bool near_diag = false, far_diag = false;
  for (int i = 0; i < 3; ++i) {
    if (ray.dir[i] == R(0)) {
      if (ray.origin[i] < bb.mins[i] || ray.origin[i] > bb.maxs[i]) return false;
    } else {
      R denom = R(1) / ray.dir[i];
      bool flip;
      R tn = (bb.mins[i] - ray.origin[i]) * denom;
      R tf = (bb.maxs[i] - ray.origin[i]) * denom;
      if (tn > tf) { flip = true; std::swap(tn, tf); } else flip = false;
      if (tn > tmin) { tmin = tn; near_side = flip ? -(i + 1) : i + 1; near_diag = false; }
      else if (tn == tmin) near_diag = true;
      if (tf < tmax) { tmax = tf; far_side = !flip ? -(i + 1) : i + 1; far_diag = false; }
      else if (tf == tmax) far_diag = true;
      if (tmax < R(0) || tmin > tmax) return false;
    }
  }
  V3<R> nn(0, 0, 0), nf(0, 0, 0);
  if (near_diag) nn = -normalize(ray.dir);
  else if (near_side < 0) nn[-near_side - 1] = R(1);
  else nn[near_side - 1] = R(-1);
  if (far_diag) nf = -normalize(ray.dir);
  else if (far_side < 0) nf[-far_side - 1] = R(-1);
  else nf[far_side - 1] = R(1);
  if (tmin < R(0)) {
    if (solid) { *toi = R(0); *normal = V3<R>(0, 0, 0); *side = far_side; return true; }
    if (tmax <= max_toi) { *toi = tmax; *normal = nf; *side = far_side; return true; }
    return false;
  }
  if (tmin <= max_toi) { *toi = tmin; *normal = nn; *side = near_side; return true; }
  return false;
}

// ray_intersection_with_support_map_with_params (ray_support_map.rs:15-62), shape at identity.
// *iters (if given) counts the support evaluations of both casts.
template <class R> inline bool ray_support_map(const Shape<R>& g, const Ray<R>& ray, R max_toi, bool solid, R* toi,
                                               V3<R>* normal, int* iters = nullptr) {
  Iso<R> id;
  id.t = V3<R>(0, 0, 0);
  id.q = Quat<R>{R(0), R(0), R(0), R(1)};
  Shape<R> origin;
  origin.type = ORIGIN; origin.a = origin.b = origin.c = R(0); origin.pts = nullptr; origin.npts = 0;
  VoronoiSimplex<R> simplex;  // the reset with the -dir support point is overwritten by cast_ray
  R t;
  V3<R> n;
  int it1 = 0, it2 = 0;
  if (iters) *iters = 0;
  const bool h1 = minkowski_ray_cast(id, smap(g), id, smap(origin), ray, max_toi, simplex, &t, &n, &it1);
  if (iters) *iters = it1;
  if (!h1) return false;
  if (solid || t != R(0)) { *toi = t; *normal = n; return true; }
  // the ray starts inside the shape: cast back from a point past the far side
  V3<R> ndir = normalize(ray.dir);
  V3<R> supp = support_point(g, id, ndir);
  R eps = R(0.001);
  R shift = dot(supp - ray.origin, ndir) + eps;
  Ray<R> nr{ray.origin + ndir * shift, -ray.dir};
  VoronoiSimplex<R> s2;
  const bool h2 = minkowski_ray_cast(id, smap(g), id, smap(origin), nr, shift + eps, s2, &t, &n, &it2);
  if (iters) *iters = it1 + it2;
  if (!h2) return false;
  R tt = shift - t;
  if (!(tt <= max_toi)) return false;
  *toi = tt;
  *normal = n;
  return true;
}

// RayCast::toi_and_normal_with_ray of Ball (ray_ball.rs:32-43, 127-142), Cuboid (ray_cuboid.rs:17-26
// -> ray_aabb.rs:45-61), Capsule and ConvexHull (ray_support_map.rs:117-163: local-frame ray,
// normal back to world).
template <class R> inline bool shape_ray_cast(const Shape<R>& g, const Iso<R>& m, const Ray<R>& ray, R max_toi,
                                              bool solid, RayHit<R>* out, int* iters = nullptr) {
  if (iters) *iters = 0;
  switch (g.type) {
    case BALL: {
      bool inside;
      R t;
      if (!ray_toi_with_ball(m.t, g.a, ray, solid, &inside, &t)) return false;
      V3<R> pos = (ray.origin + ray.dir * t) - m.t;
      V3<R> n = normalize(pos);
      if (!(t <= max_toi)) return false;
      out->toi = t;
      out->normal = inside ? -n : n;
      out->feature = 0;
      return true;
    }
    case CUBOID: {
      Ray<R> ls{inv_xform_pt(m, ray.origin), inv_xform_vec(m, ray.dir)};
      AABB<R> bb{-V3<R>(g.a, g.b, g.c), V3<R>(g.a, g.b, g.c)};
      R t;
      V3<R> n;
      int side;
      if (!ray_aabb_normal(bb, ls, max_toi, solid, &t, &n, &side)) return false;
      out->toi = t;
      out->normal = xform_vec(m, n);
      out->feature = side < 0 ? (uint32_t)(-side - 1 + 3) : (uint32_t)(side - 1);
      return true;
    }
    case CAPSULE:
    case HULL: {
      Ray<R> ls{inv_xform_pt(m, ray.origin), inv_xform_vec(m, ray.dir)};
      R t;
      V3<R> n;
      if (!ray_support_map(g, ls, max_toi, solid, &t, &n, iters)) return false;
      out->toi = t;
      out->normal = xform_vec(m, n);
      out->feature = FEAT_UNKNOWN;
      return true;
    }
    default:
      return false;
  }
}


// ================================================================ world ray queries (brute force)
// first_interference_with_ray (glue/query.rs:200-224, dbvt_broad_phase.rs:433-460) over proxies
// with stored boxes `boxes`: the best-first search returns the least-toi shape hit among the
// proxies whose box the ray enters (AABB::toi_with_ray, solid) no later than the best so far; that
// is the least toi over every proxy whose box the ray enters within max_toi.  Ties: least handle.
template <class R> inline bool world_first_ray(const std::vector<Shape<R>>& shapes, const std::vector<Iso<R>>& poses,
                                               const AABB<R>* boxes, const Ray<R>& ray, R max_toi, uint32_t* handle,
                                               RayHit<R>* hit) {
  bool found = false;
  for (size_t h = 0; h < shapes.size(); ++h) {
    R bt;
    if (!ray_aabb_toi(boxes[h], ray, max_toi, true, &bt)) continue;
    RayHit<R> r;
    if (!shape_ray_cast(shapes[h], poses[h], ray, max_toi, true, &r)) continue;
    if (!found || r.toi < hit->toi) {  // ascending handles: the first of equal tois is the least handle
      *hit = r;
      *handle = (uint32_t)h;
      found = true;
    }
  }
  return found;
}


// ================================================================ persistent narrow phase
// ConvexPolyhedronConvexPolyhedronManifoldGenerator::generate_contacts, its GJK part
// (pipeline/narrow_phase/contact_generator/convex_polyhedron_convex_polyhedron_manifold_generator.rs:
// 96-145): contact_support_map_support_map_with_params (contact_support_map_support_map.rs:38-76)
// from last_gjk_dir (None when `last_dir` is null), then last_gjk_dir <- the result's direction:
// ClosestPoints(.., dir) (EPA's normal for a penetrating pair) and NoIntersection(dir) (GJK's
// separating direction; +x when EPA fails or GJK stops at its iteration cap) set it; nothing else
// changes it (*dir_set = false).  The contact is Contact::new_wo_depth of ClosestPoints.
template <class R> inline bool polyhedra_gjk_step(const Iso<R>& m1, const Shape<R>& g1, const Iso<R>& m2,
                                                  const Shape<R>& g2, R prediction, const V3<R>* last_dir,
                                                  Contact<R>* out, V3<R>* dir_out, bool* dir_set) {
  V3<R> dir;
  if (last_dir) dir = *last_dir;
  else if (!try_normalize(m2.t - m1.t, Eps<R>::eps(), &dir)) dir = V3<R>(1, 0, 0);
  VoronoiSimplex<R> simplex;
  simplex.reset(cso_from_shapes(m1, g1, m2, g2, dir));
  GJKResult<R> r = gjk_closest_points(m1, g1, m2, g2, prediction, true, simplex);
  *dir_set = false;
  V3<R> w1, w2, n;
  if (r.status == GJK_CLOSEST) {
    w1 = r.p1; w2 = r.p2; n = r.dir;
  } else if (r.status == GJK_INTERSECTION) {
    EPA<R> epa;
    if (!epa.closest_points(m1, g1, m2, g2, simplex, &w1, &w2, &n)) {
      *dir_out = V3<R>(1, 0, 0);
      *dir_set = true;
      return false;
    }
  } else {
    if (r.status == GJK_NO_INTERSECTION) {
      *dir_out = r.dir;
      *dir_set = true;
    }
    return false;
  }
  *dir_out = n;
  *dir_set = true;
  out->world1 = w1; out->world2 = w2; out->normal = n;
  out->depth = -dot(n, w2 - w1);
  return true;
}


// ================================================================ composite shape (TriMesh) against a convex shape
// nalgebra Quaternion product, coordinates (i, j, k, w).
template <class R> inline Quat<R> quat_mul(const Quat<R>& a, const Quat<R>& b) {
  Quat<R> r;
  r.w = ((a.w * b.w - a.i * b.i) - a.j * b.j) - a.k * b.k;
  r.i = ((a.w * b.i + a.i * b.w) + a.j * b.k) - a.k * b.j;
  r.j = ((a.w * b.j - a.i * b.k) + a.j * b.w) + a.k * b.i;
  r.k = ((a.w * b.k + a.i * b.j) - a.j * b.i) + a.k * b.w;
  return r;
}
// m1.inverse() * m2 (Isometry::inverse, Isometry * Isometry).
template <class R> inline Iso<R> iso_inv_mul(const Iso<R>& m1, const Iso<R>& m2) {
  Quat<R> qi{-m1.q.i, -m1.q.j, -m1.q.k, m1.q.w};
  Iso<R> r;
  V3<R> t_inv = qrot(qi, -m1.t);
  r.t = t_inv + qrot(qi, m2.t);
  r.q = quat_mul(qi, m2.q);
  return r;
}

template <class R> inline AABB<R> triangle_aabb(const V3<R>* t) {  // local AABB: the vertices' min and max
  return AABB<R>{vmin(vmin(t[0], t[1]), t[2]), vmax(vmax(t[0], t[1]), t[2])};
}

// contact_composite_shape_shape (query/contact/contact_composite_shape_shape.rs:9-45): the
// triangles whose box intersects g2's box in the mesh frame loosened by the prediction
// (BoundingVolumeInterferencesCollector over the BVT), query::contact on each, the deepest kept
// (replace when deeper; triangles in ascending id order, so the least id among equally deep).
// tri: 3 vertices per triangle (mesh frame).
template <class R> inline bool contact_mesh_shape(const Iso<R>& m1, const V3<R>* tri, size_t ntri, const Iso<R>& m2,
                                                  const Shape<R>& g2, R prediction, Contact<R>* out,
                                                  uint32_t* face) {
  Iso<R> ls = iso_inv_mul(m1, m2);
  AABB<R> box = aabb_loosened(shape_aabb(g2, ls), prediction);
  bool found = false;
  for (size_t k = 0; k < ntri; ++k) {
    if (!aabb_intersects(triangle_aabb(tri + 3 * k), box)) continue;
    Shape<R> t;
    t.type = TRIANGLE; t.a = t.b = t.c = R(0); t.pts = tri + 3 * k; t.npts = 3;
    Contact<R> c;
    c.depth = R(0);
    if (!contact(m1, t, m2, g2, prediction, &c)) continue;
    if (!found || c.depth > out->depth) {
      *out = c;
      *face = (uint32_t)k;
      found = true;
    }
  }
  return found;
}

// distance_composite_shape_shape (query/distance/distance_composite_shape_shape.rs:9-101): the
// best-first search over the BVT returns the least query::distance among the triangles whose
// Minkowski-sum box distance (the visitor's cost, :70-79) is below the best so far; visiting the
// leaves by increasing cost and stopping at the first cost not below the best gives that value.
template <class R> inline R distance_mesh_shape(const Iso<R>& m1, const V3<R>* tri, size_t ntri, const Iso<R>& m2,
                                                const Shape<R>& g2) {
  Iso<R> ls = iso_inv_mul(m1, m2);
  AABB<R> b2 = shape_aabb(g2, ls);
  V3<R> shift = -aabb_center(b2);
  V3<R> margin = (b2.maxs - b2.mins) * R(0.5);
  std::vector<std::pair<R, size_t>> order(ntri);
  for (size_t k = 0; k < ntri; ++k) {
    AABB<R> t = triangle_aabb(tri + 3 * k);
    V3<R> lo = (t.mins + shift) + (-margin), hi = (t.maxs + shift) + margin;
    V3<R> z(0, 0, 0);
    V3<R> mins_pt = lo - z, pt_maxs = z - hi;
    V3<R> sh(std::fmax(std::fmax(mins_pt.x, pt_maxs.x), R(0)), std::fmax(std::fmax(mins_pt.y, pt_maxs.y), R(0)),
             std::fmax(std::fmax(mins_pt.z, pt_maxs.z), R(0)));
    order[k] = {norm(sh), k};
  }
  std::sort(order.begin(), order.end());
  R best = Eps<R>::maxv();
  for (auto& e : order) {
    if (!(e.first < best)) break;
    Shape<R> t;
    t.type = TRIANGLE; t.a = t.b = t.c = R(0); t.pts = tri + 3 * e.second; t.npts = 3;
    R d = distance(m1, t, m2, g2);
    if (d < best) best = d;
  }
  return best;
}


// ================================================================ contact manifolds
// The narrow phase's contact-manifold generators for Ball, Cuboid, Capsule and ConvexHull
// (pipeline/narrow_phase/contact_generator/*), the polygonal-feature clipping they use
// (shape/convex_polygonal_feature3.rs) and ContactManifold's distance-based contact tracking
// (query/contact/contact_manifold.rs).
//
// Feature ids: kind in the top two bits (0 vertex, 1 edge, 2 face, 3 unknown), id below.
constexpr uint32_t FEAT_V = 0u, FEAT_E = 1u, FEAT_F = 2u;
constexpr uint32_t FEAT_NONE = 0xC0000000u;  // FeatureId::Unknown
inline uint32_t feat(uint32_t kind, uint32_t id) { return (kind << 30) | id; }
inline uint32_t feat_kind(uint32_t f) { return f >> 30; }
inline uint32_t feat_id(uint32_t f) { return f & 0x3FFFFFFFu; }

// ConvexHull::try_new (shape/convex.rs:107-324): polygonal faces from the hull's outward
// triangles.  An edge between triangles whose normals' dot exceeds 1 - sqrt(eps) is deleted; a
// face is the contour walked around deleted edges, with its first triangle's normal.
template <class R> struct HullPolyO {
  std::vector<V3<R>> fn;                 // face normals
  std::vector<uint32_t> ffirst, fnum;    // face loops in lvert / ledge
  std::vector<uint32_t> lvert, ledge;    // vertices_adj_to_face, edges_adj_to_face
  std::vector<uint32_t> ev0, ev1;        // edge end vertices
  std::vector<V3<R>> edir;               // edge unit directions
  std::vector<uint32_t> vfirst, vnum;    // per vertex: adjacency run
  std::vector<uint32_t> vface, vedge;    // faces_adj_to_vertex, edges_adj_to_vertex
};

template <class R> inline bool hull_poly_build_o(const V3<R>* pts, uint32_t npts, const uint32_t* tris, uint32_t ntri,
                                                 HullPolyO<R>& H) {
  H = HullPolyO<R>();
  const R eps = std::sqrt(Eps<R>::eps());
  struct E { uint32_t v0, v1, f0, f1; V3<R> dir; bool del; };
  struct T { uint32_t v[3], e[3]; V3<R> n; int64_t parent; };
  std::vector<E> edges;
  std::vector<T> tri(ntri);
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> emap(npts);  // SortedPair -> edge
  for (uint32_t t = 0; t < ntri; ++t) {
    const uint32_t* vtx = tris + 3 * (size_t)t;
    for (int i = 0; i < 3; ++i)
      if (vtx[i] >= npts) return false;
    for (int i1 = 0; i1 < 3; ++i1) {
      int i2 = (i1 + 1) % 3;
      uint32_t lo = std::min(vtx[i1], vtx[i2]), hi = std::max(vtx[i1], vtx[i2]);
      int64_t found = -1;
      for (auto& e : emap[lo]) if (e.first == hi) found = e.second;
      if (found >= 0) {
        tri[t].e[i1] = (uint32_t)found;
        edges[(size_t)found].f1 = t;
      } else {
        V3<R> d;
        if (!try_normalize(pts[vtx[i2]] - pts[vtx[i1]], Eps<R>::eps(), &d)) return false;
        tri[t].e[i1] = (uint32_t)edges.size();
        emap[lo].push_back({hi, (uint32_t)edges.size()});
        edges.push_back(E{vtx[i1], vtx[i2], t, 0, d, false});
      }
    }
    V3<R> n;  // utils::ccw_face_normal (utils/ccw_face_normal.rs:21-27)
    if (!try_normalize(cross(pts[vtx[1]] - pts[vtx[0]], pts[vtx[2]] - pts[vtx[0]]), Eps<R>::eps(), &n)) return false;
    for (int k = 0; k < 3; ++k) tri[t].v[k] = vtx[k];
    tri[t].n = n;
    tri[t].parent = -1;
  }
  int64_t valid_edges = 0;
  for (auto& e : edges) {
    if (dot(tri[e.f0].n, tri[e.f1].n) > R(1) - eps) e.del = true;
    else valid_edges++;
  }
  std::vector<uint32_t> face_tri;
  for (uint32_t i = 0; i < ntri; ++i) {
    if (tri[i].parent >= 0) continue;
    for (int j1 = 0; j1 < 3; ++j1) {
      if (edges[tri[i].e[j1]].del) continue;
      uint32_t face_id = (uint32_t)H.ffirst.size(), first = (uint32_t)H.ledge.size(), num = 1;
      H.ledge.push_back(tri[i].e[j1]);
      H.lvert.push_back(tri[i].v[j1]);
      uint32_t start = tri[i].v[j1], ct = i;
      int ce = (j1 + 1) % 3;
      size_t steps = 0;
      while (tri[ct].v[ce] != start) {
        if (++steps > 8 * (size_t)ntri + 16) return false;
        uint32_t cur_e = tri[ct].e[ce], cur_v = tri[ct].v[ce];
        tri[ct].parent = face_id;
        if (!edges[cur_e].del) {
          H.ledge.push_back(cur_e);
          H.lvert.push_back(cur_v);
          num++;
          ce = (ce + 1) % 3;
        } else {
          ct = ct == edges[cur_e].f0 ? edges[cur_e].f1 : edges[cur_e].f0;  // Edge::other_triangle
          int nx = -1;
          for (int k = 0; k < 3; ++k) if (tri[ct].e[k] == cur_e) nx = (k + 1) % 3;  // Triangle::next_edge_id
          if (nx < 0 || tri[ct].v[nx] != cur_v) return false;
          ce = nx;
        }
      }
      if (num > 2) { H.ffirst.push_back(first); H.fnum.push_back(num); face_tri.push_back(i); }
      break;
    }
  }
  for (uint32_t t : face_tri) H.fn.push_back(tri[t].n);
  for (auto& e : edges) { H.ev0.push_back(e.v0); H.ev1.push_back(e.v1); H.edir.push_back(e.dir); }
  H.vfirst.assign(npts, 0);
  H.vnum.assign(npts, 0);
  for (size_t f = 0; f < H.ffirst.size(); ++f)
    for (uint32_t k = H.ffirst[f]; k < H.ffirst[f] + H.fnum[f]; ++k) H.vnum[H.lvert[k]]++;
  uint32_t total = 0;
  for (uint32_t v = 0; v < npts; ++v) { H.vfirst[v] = total; total += H.vnum[v]; H.vnum[v] = 0; }
  H.vface.assign(total, 0);
  H.vedge.assign(total, 0);
  for (size_t f = 0; f < H.ffirst.size(); ++f)
    for (uint32_t k = H.ffirst[f]; k < H.ffirst[f] + H.fnum[f]; ++k) {
      uint32_t v = H.lvert[k];
      H.vface[H.vfirst[v] + H.vnum[v]] = (uint32_t)f;
      H.vedge[H.vfirst[v] + H.vnum[v]] = H.ledge[k];
      H.vnum[v]++;
    }
  int64_t valid_vertices = 0;
  for (uint32_t v = 0; v < npts; ++v) valid_vertices += H.vnum[v] != 0;
  return valid_vertices + (int64_t)H.ffirst.size() - valid_edges == 2;  // Euler's characteristic
}

// ConvexPolygonalFeature (shape/convex_polygonal_feature3.rs:38-56); edge normals are left out
// (only ConvexPolygonalFeature::project_point reads them, and no generator here calls it).
template <class R> struct PolyFeat {
  std::vector<V3<R>> v;
  std::vector<uint32_t> vid, eid;
  bool has_n = false;
  V3<R> n;
  uint32_t fid = FEAT_NONE;
  void clear() { v.clear(); vid.clear(); eid.clear(); has_n = false; fid = FEAT_NONE; }
  void push(V3<R> p, uint32_t id) { v.push_back(p); vid.push_back(id); }
  size_t nedges() const { return v.size() == 1 ? 0 : (v.size() == 2 ? 1 : v.size()); }
  void transform_by(const Iso<R>& m) {
    for (auto& p : v) p = xform_pt(m, p);
    if (has_n) n = xform_vec(m, n);
  }
  // ConvexPolygonalFeature::edge: the segment starting at the vertex whose edge id matches
  bool edge(uint32_t id, V3<R>* a, V3<R>* b) const {
    for (size_t i1 = 0; i1 < v.size() && i1 < eid.size(); ++i1)
      if (eid[i1] == id) { *a = v[i1]; *b = v[(i1 + 1) % v.size()]; return true; }
    return false;
  }
};

// Cuboid::face (shape/cuboid.rs:185-276), 3D.
template <class R> inline void cuboid_face(V3<R> h, uint32_t i, PolyFeat<R>& out) {
  out.clear();
  uint32_t i1;
  R sign;
  if (i < 3) { i1 = i; sign = R(1); } else { i1 = i - 3; sign = R(-1); }
  uint32_t i2 = (i1 + 1) % 3, i3 = (i1 + 2) % 3;
  uint32_t ei2 = sign > R(0) ? i2 : i3, ei3 = sign > R(0) ? i3 : i2;
  uint32_t mask_i2 = ~(1u << ei2), mask_i3 = ~(1u << ei3);
  V3<R> vertex = h;
  vertex[i1] = vertex[i1] * sign;
  uint32_t sbit = sign < R(0) ? 1u : 0u, msbit = sign < R(0) ? 0u : 1u;
  uint32_t vertex_id = sbit << i1;
  out.push(vertex, feat(FEAT_V, vertex_id));
  out.eid.push_back(feat(FEAT_E, ei2 | ((vertex_id & mask_i2) << 2)));
  vertex[i2] = -sign * h[i2];
  vertex[i3] = sign * h[i3];
  vertex_id |= (msbit << i2) | (sbit << i3);
  out.push(vertex, feat(FEAT_V, vertex_id));
  out.eid.push_back(feat(FEAT_E, ei3 | ((vertex_id & mask_i3) << 2)));
  vertex[i2] = -h[i2];
  vertex[i3] = -h[i3];
  vertex_id |= (1u << i2) | (1u << i3);
  out.push(vertex, feat(FEAT_V, vertex_id));
  out.eid.push_back(feat(FEAT_E, ei2 | ((vertex_id & mask_i2) << 2)));
  vertex[i2] = sign * h[i2];
  vertex[i3] = -sign * h[i3];
  vertex_id = (sbit << i1) | (sbit << i2) | (msbit << i3);
  out.push(vertex, feat(FEAT_V, vertex_id));
  out.eid.push_back(feat(FEAT_E, ei3 | ((vertex_id & mask_i3) << 2)));
  V3<R> nn(0, 0, 0);
  nn[i1] = sign;
  out.has_n = true;
  out.n = nn;
  out.fid = feat(FEAT_F, sign > R(0) ? i1 : i1 + 3);
}

// ConvexHull::face (shape/convex.rs:443-460).
template <class R> inline void hull_face(const Shape<R>& g, uint32_t f, PolyFeat<R>& out) {
  out.clear();
  const HullPolyO<R>& H = *g.poly;
  for (uint32_t k = H.ffirst[f]; k < H.ffirst[f] + H.fnum[f]; ++k) {
    out.push(g.pts[H.lvert[k]], feat(FEAT_V, H.lvert[k]));
    out.eid.push_back(feat(FEAT_E, H.ledge[k]));
  }
  out.has_n = true;
  out.n = H.fn[f];
  out.fid = feat(FEAT_F, f);
}

// R::signum (Rust): +1 for +0 and positives, -1 for -0 and negatives, NaN for NaN.
template <class R> inline R rsignum(R x) { return std::isnan(x) ? x : (std::signbit(x) ? R(-1) : R(1)); }

// ConvexPolyhedron::support_face_toward for Cuboid (cuboid.rs:279-307), ConvexHull
// (convex.rs:487-509) and Segment (segment.rs:291-304).
template <class R> inline void support_face_toward(const Shape<R>& g, const Iso<R>& m, V3<R> dir, PolyFeat<R>& out) {
  out.clear();
  if (g.type == CUBOID) {
    V3<R> ld = inv_xform_vec(m, dir);
    int iamax = 0;
    R amax = std::fabs(ld[0]);
    for (int i = 1; i < 3; ++i) if (std::fabs(ld[i]) > amax) { amax = std::fabs(ld[i]); iamax = i; }
    cuboid_face(V3<R>(g.a, g.b, g.c), ld[iamax] > R(0) ? (uint32_t)iamax : (uint32_t)iamax + 3, out);
    out.transform_by(m);
  } else if (g.type == HULL) {
    V3<R> ld = inv_xform_vec(m, dir);
    const HullPolyO<R>& H = *g.poly;
    uint32_t best = 0;
    R max_dot = dot(H.fn[0], ld);
    for (uint32_t f = 1; f < H.fn.size(); ++f) {
      R d = dot(H.fn[f], ld);
      if (d > max_dot) { max_dot = d; best = f; }
    }
    hull_face(g, best, out);
    out.transform_by(m);
  } else {  // SEGMENT
    out.push(V3<R>(R(0), -g.a, R(0)), feat(FEAT_V, 0));
    out.push(V3<R>(R(0), g.a, R(0)), feat(FEAT_V, 1));
    out.eid.push_back(feat(FEAT_E, 0));
    out.fid = feat(FEAT_E, 0);
    out.transform_by(m);
  }
}

// ConvexHull::support_feature_id_toward_eps (convex.rs:388-415): sin and cos of the angle given.
template <class R> inline uint32_t hull_support_feature_id(const Shape<R>& g, V3<R> ld, R seps, R ceps) {
  const HullPolyO<R>& H = *g.poly;
  uint32_t sv = point_cloud_support_id(ld, g.pts, g.npts);
  for (uint32_t i = 0; i < H.vnum[sv]; ++i) {
    uint32_t f = H.vface[H.vfirst[sv] + i];
    if (dot(H.fn[f], ld) >= ceps) return feat(FEAT_F, f);
  }
  for (uint32_t i = 0; i < H.vnum[sv]; ++i) {
    uint32_t e = H.vedge[H.vfirst[sv] + i];
    if (std::fabs(dot(H.edir[e], ld)) <= seps) return feat(FEAT_E, e);
  }
  return feat(FEAT_V, sv);
}

// ConvexPolyhedron::support_feature_toward for Cuboid (cuboid.rs:309-405), ConvexHull
// (convex.rs:511-540) and Segment (segment.rs:306-344); cang/sang: cos and sin of the angle.
template <class R> inline void support_feature_toward(const Shape<R>& g, const Iso<R>& m, V3<R> dir, R cang, R sang,
                                                      PolyFeat<R>& out) {
  out.clear();
  if (g.type == CUBOID) {
    V3<R> h(g.a, g.b, g.c);
    V3<R> ld = inv_xform_vec(m, dir);
    V3<R> sp = h;
    uint32_t spid = 0;
    for (int i1 = 0; i1 < 3; ++i1) {
      R sign = rsignum(ld[i1]);
      if (sign * ld[i1] >= cang) {
        cuboid_face(h, sign > R(0) ? (uint32_t)i1 : (uint32_t)i1 + 3, out);
        out.transform_by(m);
        return;
      } else if (sign < R(0)) {
        sp[i1] = sp[i1] * sign;
        spid |= 1u << i1;
      }
    }
    for (int i = 0; i < 3; ++i) {
      R sign = rsignum(ld[i]);
      if (sign * ld[i] <= sang) {
        sp[i] = -h[i];
        V3<R> p1 = sp;
        sp[i] = h[i];
        V3<R> p2 = sp;
        uint32_t p2id = spid & ~(1u << i);
        out.push(xform_pt(m, p1), feat(FEAT_V, spid | (1u << i)));
        out.push(xform_pt(m, p2), feat(FEAT_V, p2id));
        uint32_t e = feat(FEAT_E, (uint32_t)i | (p2id << 2));
        out.eid.push_back(e);
        out.fid = e;
        return;
      }
    }
    out.push(xform_pt(m, sp), feat(FEAT_V, spid));
    out.fid = feat(FEAT_V, spid);
  } else if (g.type == HULL) {
    V3<R> ld = inv_xform_vec(m, dir);
    uint32_t f = hull_support_feature_id(g, ld, sang, cang);
    const HullPolyO<R>& H = *g.poly;
    if (feat_kind(f) == FEAT_V) {
      out.push(g.pts[feat_id(f)], f);
      out.fid = f;
    } else if (feat_kind(f) == FEAT_E) {
      uint32_t e = feat_id(f);
      out.push(g.pts[H.ev0[e]], feat(FEAT_V, H.ev0[e]));
      out.push(g.pts[H.ev1[e]], feat(FEAT_V, H.ev1[e]));
      out.fid = f;
      out.eid.push_back(f);
    } else {
      hull_face(g, feat_id(f), out);
    }
    out.transform_by(m);
  } else {  // SEGMENT: the segment placed by m
    V3<R> a = xform_pt(m, V3<R>(R(0), -g.a, R(0))), b = xform_pt(m, V3<R>(R(0), g.a, R(0)));
    V3<R> sd;
    if (!try_normalize(b - a, Eps<R>::eps(), &sd)) return;
    R c = dot(dir, sd);
    if (c > sang) { out.fid = feat(FEAT_V, 1); out.push(b, feat(FEAT_V, 1)); }
    else if (c < -sang) { out.fid = feat(FEAT_V, 0); out.push(a, feat(FEAT_V, 0)); }
    else {
      out.push(a, feat(FEAT_V, 0));
      out.push(b, feat(FEAT_V, 1));
      out.eid.push_back(feat(FEAT_E, 0));
      out.fid = feat(FEAT_E, 0);
    }
  }
}

// 2D helpers of the clipping
template <class R> struct P2 { R x, y; };
// utils::point_in_poly2d (utils/point_in_poly2d.rs:5-26)
template <class R> inline bool point_in_poly2d(P2<R> pt, const std::vector<P2<R>>& poly) {
  if (poly.empty()) return false;
  R sign = R(0);
  for (size_t i1 = 0; i1 < poly.size(); ++i1) {
    size_t i2 = (i1 + 1) % poly.size();
    R sx = poly[i2].x - poly[i1].x, sy = poly[i2].y - poly[i1].y;
    R dx = pt.x - poly[i1].x, dy = pt.y - poly[i1].y;
    R perp = dx * sy - dy * sx;
    if (sign == R(0)) sign = perp;
    else if (sign * perp < R(0)) return false;
  }
  return true;
}

// approx::ulps_eq!(a, b) with the default epsilon (EPSILON) and max_ulps (4).
inline bool ulps_eq_r(float a, float b) {
  if ((a > b ? a - b : b - a) <= Eps<float>::eps()) return true;
  if (std::signbit(a) != std::signbit(b) || std::isnan(a) || std::isnan(b)) return false;
  uint32_t ia, ib;
  std::memcpy(&ia, &a, 4);
  std::memcpy(&ib, &b, 4);
  return (ia <= ib ? ib - ia : ia - ib) <= 4u;
}
inline bool ulps_eq_r(double a, double b) {
  if ((a > b ? a - b : b - a) <= Eps<double>::eps()) return true;
  if (std::signbit(a) != std::signbit(b) || std::isnan(a) || std::isnan(b)) return false;
  uint64_t ia, ib;
  std::memcpy(&ia, &a, 8);
  std::memcpy(&ib, &b, 8);
  return (ia <= ib ? ib - ia : ia - ib) <= 4u;
}
// nalgebra::clamp
template <class R> inline R na_clamp(R v, R lo, R hi) { return v > lo ? (v < hi ? v : hi) : lo; }

// closest_points_segment_segment_with_locations_nD (query/closest_points/
// closest_points_segment_segment.rs:44-141) in 2D: true when both points lie inside their
// segments (SegmentPointLocation::OnEdge), with the parameters s and t.
template <class R> inline bool seg_seg_2d_on_edges(P2<R> p1, P2<R> q1, P2<R> p2, P2<R> q2, R* s_out, R* t_out) {
  const R eps = Eps<R>::eps();
  R d1x = q1.x - p1.x, d1y = q1.y - p1.y, d2x = q2.x - p2.x, d2y = q2.y - p2.y;
  R rx = p1.x - p2.x, ry = p1.y - p2.y;
  R a = d1x * d1x + d1y * d1y, e = d2x * d2x + d2y * d2y, f = d2x * rx + d2y * ry;
  R s, t;
  if (a <= eps && e <= eps) { s = R(0); t = R(0); }
  else if (a <= eps) { s = R(0); t = na_clamp(f / e, R(0), R(1)); }
  else {
    R c = d1x * rx + d1y * ry;
    if (e <= eps) { t = R(0); s = na_clamp(-c / a, R(0), R(1)); }
    else {
      R b = d1x * d2x + d1y * d2y;
      R ae = a * e, bb = b * b;
      R denom = ae - bb;
      bool parallel = denom <= eps || ulps_eq_r(ae, bb);
      s = !parallel ? na_clamp((b * f - c * e) / denom, R(0), R(1)) : R(0);
      t = (b * s + f) / e;
      if (t < R(0)) { t = R(0); s = na_clamp(-c / a, R(0), R(1)); }
      else if (t > R(1)) { t = R(1); s = na_clamp((b - c) / a, R(0), R(1)); }
    }
  }
  *s_out = s;
  *t_out = t;
  return s != R(0) && s != R(1) && t != R(0) && t != R(1);
}

template <class R> struct ClipContact { Contact<R> c; uint32_t f1, f2; };

// ConvexPolygonalFeature::clip (convex_polygonal_feature3.rs:217-352).
template <class R> inline void clip_features(const PolyFeat<R>& s, const PolyFeat<R>& o, V3<R> normal, R linear,
                                             std::vector<ClipContact<R>>& out) {
  if (s.v.size() <= 2 && o.v.size() <= 2) return;
  // Vector::orthonormal_subspace_basis(&[normal]): a x normal, then a
  V3<R> a;
  if (std::fabs(normal.x) > std::fabs(normal.y)) a = V3<R>(normal.z, R(0), -normal.x);
  else a = V3<R>(R(0), -normal.z, normal.y);
  a = normalize(a);
  V3<R> b0 = cross(a, normal), b1 = a;
  V3<R> ref = s.v[0];
  std::vector<P2<R>> p1, p2;
  for (auto& p : s.v) { V3<R> d = p - ref; p1.push_back(P2<R>{dot(b0, d), dot(b1, d)}); }
  for (auto& p : o.v) { V3<R> d = p - ref; p2.push_back(P2<R>{dot(b0, d), dot(b1, d)}); }
  auto emit = [&](V3<R> w1, V3<R> w2, uint32_t f1, uint32_t f2) {
    Contact<R> c;  // Contact::new_wo_depth (contact.rs:43-46)
    c.world1 = w1; c.world2 = w2; c.normal = normal;
    c.depth = -dot(normal, w2 - w1);
    if (-c.depth <= linear) out.push_back(ClipContact<R>{c, f1, f2});
  };
  // query::line_toi_with_plane (query/ray/ray_plane.rs:9-23)
  auto line_toi = [&](V3<R> pc, V3<R> pn, V3<R> lo, V3<R> ld, R* toi) {
    V3<R> dpos = pc - lo;
    R denom = dot(pn, ld);
    if (relative_eq_r(denom, R(0))) return false;
    *toi = dot(pn, dpos) / denom;
    return true;
  };
  if (p2.size() > 2) {
    for (size_t i = 0; i < p1.size(); ++i) {
      if (!point_in_poly2d(p1[i], p2)) continue;
      V3<R> origin = (ref + b0 * p1[i].x) + b1 * p1[i].y;
      R toi;
      if (line_toi(o.v[0], o.n, origin, normal, &toi)) emit(s.v[i], origin + normal * toi, s.vid[i], o.fid);
    }
  }
  if (p1.size() > 2) {
    for (size_t i = 0; i < p2.size(); ++i) {
      if (!point_in_poly2d(p2[i], p1)) continue;
      V3<R> origin = (ref + b0 * p2[i].x) + b1 * p2[i].y;
      R toi;
      if (line_toi(s.v[0], s.n, origin, normal, &toi)) emit(origin + normal * toi, o.v[i], s.fid, o.vid[i]);
    }
  }
  size_t ne1 = s.nedges(), ne2 = o.nedges();
  for (size_t i1 = 0; i1 < ne1; ++i1) {
    size_t j1 = (i1 + 1) % p1.size();
    for (size_t i2 = 0; i2 < ne2; ++i2) {
      size_t j2 = (i2 + 1) % p2.size();
      R sp, tp;
      if (!seg_seg_2d_on_edges(p1[i1], p1[j1], p2[i2], p2[j2], &sp, &tp)) continue;
      // Segment::point_at(OnEdge([1 - s, s])) (shape/segment.rs:128-137)
      V3<R> w1 = s.v[i1] * (R(1) - sp) + s.v[j1] * sp;
      V3<R> w2 = o.v[i2] * (R(1) - tp) + o.v[j2] * tp;
      emit(w1, w2, s.eid[i1], o.eid[i2]);
    }
  }
}

// ContactManifold with ContactTrackingMode::DistanceBased(0.02) (contact_manifold.rs:44-277):
// the cache holds one tracking point per contact; save_cache_and_clear keeps the entries whose
// contact was in the last update and marks them old.
template <class R> struct TrackedEntry {
  V3<R> track;
  bool active;
  Contact<R> c;
  uint32_t f1, f2;
};
template <class R> struct ManifoldState {
  bool has_dir = false;  // the convex-polyhedron generator's last_gjk_dir
  V3<R> dir;
  std::vector<V3<R>> cache;  // tracking points of the last update's contacts, in cache order
};

// A generator's contact preprocessor chain: the capsule's (capsule.rs:86-127) when the shape
// is a capsule's segment (radius r > 0 marks it).
template <class R> inline bool capsule_preprocess(R r, bool is_first, Contact<R>& c, uint32_t& f1, uint32_t& f2) {
  uint32_t& f = is_first ? f1 : f2;
  uint32_t k = feat_kind(f);
  if (k == FEAT_V) f = feat(FEAT_F, feat_id(f));
  else if (k == FEAT_E) f = feat(FEAT_F, 2);
  else if (k == FEAT_F) f = feat(FEAT_F, 2 + feat_id(f));
  else return false;
  if (is_first) c.world1 = c.world1 + c.normal * r;
  else c.world2 = c.world2 - c.normal * r;
  c.depth = c.depth + r;
  return true;
}

template <class R> inline void manifold_push(std::vector<TrackedEntry<R>>& cache, Contact<R> c, uint32_t f1, uint32_t f2,
                                             V3<R> track, R cap1, R cap2) {
  if (cap1 > R(0) && !capsule_preprocess(cap1, true, c, f1, f2)) return;
  if (cap2 > R(0) && !capsule_preprocess(cap2, false, c, f1, f2)) return;
  const R threshold = R(0.02);
  size_t closest = cache.size();
  R closest_dist = threshold * threshold;
  for (size_t i = 0; i < cache.size(); ++i) {
    R d = norm_sq(cache[i].track - track);
    if (d < closest_dist) { closest_dist = d; closest = i; }
  }
  if (closest == cache.size()) {
    cache.push_back(TrackedEntry<R>{track, true, c, f1, f2});
    return;
  }
  TrackedEntry<R>& e = cache[closest];
  if (e.active) {
    if (c.depth <= e.c.depth) return;  // the contact already there is deeper
  } else {
    e.active = true;
  }
  e.c = c; e.f1 = f1; e.f2 = f2;
  e.track = track;
}

// ConvexPolygonalFeature::add_contact_to_manifold (convex_polygonal_feature3.rs:355-400): the
// contact is dropped when a feature is unknown or names a degenerate edge; the tracking point is
// the contact's first point in the first shape's frame.
template <class R> inline void add_contact_to_manifold(const PolyFeat<R>& F1, const PolyFeat<R>& F2, const ClipContact<R>& cc,
                                                       const Iso<R>& m1, R cap1, R cap2,
                                                       std::vector<TrackedEntry<R>>& cache) {
  V3<R> local1 = inv_xform_pt(m1, cc.c.world1);
  for (int side = 0; side < 2; ++side) {
    uint32_t f = side == 0 ? cc.f1 : cc.f2;
    const PolyFeat<R>& F = side == 0 ? F1 : F2;
    if (f == FEAT_NONE) return;
    if (feat_kind(f) == FEAT_E) {
      V3<R> a, b, d;
      if (!F.edge(f, &a, &b)) return;
      if (!try_normalize(b - a, Eps<R>::eps(), &d)) return;
    }
  }
  manifold_push(cache, cc.c, cc.f1, cc.f2, local1, cap1, cap2);
}

// BallConvexPolyhedronManifoldGenerator (ball_convex_polyhedron_manifold_generator.rs:28-121):
// ball (center c, radius r) against a Cuboid, ConvexHull or capsule segment; flip: the ball is
// the pair's second shape.  ceps1/seps1: cos and sin of one degree (the hull feature's angle).
template <class R> inline void ball_polyhedron_generator(V3<R> c, R r, const Iso<R>& m2, const Shape<R>& g2, R linear,
                                                         bool flip, R cap_poly, std::vector<TrackedEntry<R>>& cache) {
  Contact<R> ct;
  int fkind = 3, fid = 0;
  bool inside = false;
  if (!contact_ball_convex(c, r, m2, g2, linear, &ct, &fkind, &fid, &inside)) return;
  uint32_t f2;
  if (g2.type == HULL) {
    // ConvexHull::project_point_with_feature (point_support_map.rs:96-118)
    V3<R> dpt = c - ct.world2;
    V3<R> ld = inside ? inv_xform_vec(m2, -dpt) : inv_xform_vec(m2, dpt);
    V3<R> u;
    if (!try_normalize(ld, Eps<R>::eps(), &u)) return;  // FeatureId::Unknown: the reference panics here
    const R one_degree = R(3.14159265358979323846 / 180.0);
    f2 = hull_support_feature_id(g2, u, std::sin(one_degree), std::cos(one_degree));
  } else {
    if (fkind == 3) return;  // FeatureId::Unknown: the reference panics here
    f2 = fkind == 0 ? feat(FEAT_F, (uint32_t)fid) : fkind == 1 ? feat(FEAT_E, (uint32_t)fid) : feat(FEAT_V, (uint32_t)fid);
  }
  const uint32_t ball = feat(FEAT_F, 0);
  if (!flip) manifold_push(cache, ct, ball, f2, V3<R>(0, 0, 0), R(0), cap_poly);
  else {
    Contact<R> fl;  // Contact::new(world2, world1, -normal, depth)
    fl.world1 = ct.world2; fl.world2 = ct.world1; fl.normal = -ct.normal; fl.depth = ct.depth;
    manifold_push(cache, fl, f2, ball, V3<R>(0, 0, 0), cap_poly, R(0));
  }
}

// ConvexPolyhedronConvexPolyhedronManifoldGenerator::generate_contacts
// (convex_polyhedron_convex_polyhedron_manifold_generator.rs:83-167).
template <class R> inline void polyhedra_generator(const Iso<R>& m1, const Shape<R>& g1, const Iso<R>& m2, const Shape<R>& g2,
                                                   R linear, R c1, R s1, R c2, R s2, R cap1, R cap2, ManifoldState<R>& st,
                                                   std::vector<TrackedEntry<R>>& cache) {
  Contact<R> ct;
  V3<R> nd;
  bool set = false;
  bool ok = polyhedra_gjk_step(m1, g1, m2, g2, linear, st.has_dir ? &st.dir : nullptr, &ct, &nd, &set);
  if (set) { st.has_dir = true; st.dir = nd; }
  if (!ok) return;
  PolyFeat<R> F1, F2;
  std::vector<ClipContact<R>> nc;
  if (ct.depth > R(0)) {
    support_face_toward(g1, m1, ct.normal, F1);
    support_face_toward(g2, m2, -ct.normal, F2);
  } else {
    support_feature_toward(g1, m1, ct.normal, c1, s1, F1);
    support_feature_toward(g2, m2, -ct.normal, c2, s2, F2);
  }
  clip_features(F1, F2, ct.normal, linear, nc);
  if (nc.empty()) nc.push_back(ClipContact<R>{ct, F1.fid, F2.fid});
  for (auto& cc : nc) add_contact_to_manifold(F1, F2, cc, m1, cap1, cap2, cache);
}

// One update of a pair's manifold (NarrowPhase::update_contact, narrow_phase.rs:56-80, with the
// generator DefaultContactDispatcher picks, default_contact_dispatcher.rs:27-97).  c1/s1, c2/s2:
// cos and sin of the prediction's angular1 / angular2.
template <class R> inline void manifold_update(const Iso<R>& m1, const Shape<R>& g1, const Iso<R>& m2, const Shape<R>& g2,
                                               R linear, R c1, R s1, R c2, R s2, ManifoldState<R>& st,
                                               std::vector<TrackedEntry<R>>& cache) {
  cache.clear();  // save_cache_and_clear: the last update's contacts become old entries
  for (auto& p : st.cache) cache.push_back(TrackedEntry<R>{p, false, Contact<R>{}, FEAT_NONE, FEAT_NONE});
  Shape<R> a = g1, b = g2;
  R cap1 = R(0), cap2 = R(0), lin = linear;
  // CapsuleCapsuleManifoldGenerator / CapsuleShapeManifoldGenerator: capsules become their
  // segments, the prediction grows by each radius, and each capsule adds its preprocessor
  if (g1.type == CAPSULE) { a.type = SEGMENT; cap1 = g1.b; lin = lin + g1.b; }
  if (g2.type == CAPSULE) { b.type = SEGMENT; cap2 = g2.b; lin = lin + g2.b; }
  if (a.type == BALL && b.type == BALL) {
    // BallBallManifoldGenerator (ball_ball_manifold_generator.rs:27-66)
    Contact<R> ct;
    if (contact_ball_ball(m1.t, g1.a, m2.t, g2.a, linear, &ct))
      manifold_push(cache, ct, feat(FEAT_F, 0), feat(FEAT_F, 0), V3<R>(0, 0, 0), R(0), R(0));
  } else if (a.type == BALL) {
    ball_polyhedron_generator(m1.t, g1.a, m2, b, lin, false, cap2, cache);
  } else if (b.type == BALL) {
    ball_polyhedron_generator(m2.t, g2.a, m1, a, lin, true, cap1, cache);
  } else {
    polyhedra_generator(m1, a, m2, b, lin, c1, s1, c2, s2, cap1, cap2, st, cache);
  }
  st.cache.clear();
  for (auto& e : cache) if (e.active) st.cache.push_back(e.track);
}

}  // namespace ncxo
