// ncx_gjk.h -- device support maps, CSO points, the Voronoi simplex and GJK, shared by the
// contact/distance, time-of-impact, shape ray-cast and composite-shape kernels.
//
// One group of G lanes (a power of two inside a 64-wide wavefront) evaluates one query; every
// lane runs the same control flow on the same values and the lanes split ConvexHull vertex scans.
// Arithmetic restates the reference in f32 with its operation order:
//   supports  src/shape/{ball,cuboid,capsule,convex}.rs, src/utils/point_cloud_support_point.rs
//   simplex   src/query/algorithms/voronoi_simplex3.rs:49-280
//   GJK       src/query/algorithms/gjk.rs:76-177
#pragma once
#include "ncx_internal.h"

#define SHAPE_ORIGIN 4u
#define GJK_S_INTERSECTION 0
#define GJK_S_CLOSEST 1
#define GJK_S_NOINTER 3
#define GJK_S_NAN 4
#define GJK_S_CAP 5  // iteration cap reached (NoIntersection(+x) in the reference)

struct ShapeD {
  uint32_t type, off, cnt;
  float a, b, c;
};

__device__ __forceinline__ ShapeD load_shape(const ncx_shape* __restrict__ s, uint32_t i) {
  const uint4* p = (const uint4*)(s + i);
  uint4 u = p[0];
  uint4 v = p[1];
  ShapeD g;
  g.type = u.x;
  g.off = u.y;
  g.cnt = u.z;
  g.a = __uint_as_float(u.w);
  g.b = __uint_as_float(v.x);
  g.c = __uint_as_float(v.y);
  return g;
}

struct Pool {
  const float* __restrict__ x;
  const float* __restrict__ y;
  const float* __restrict__ z;
  // TriMesh triangles (three float4 per triangle: a with the face id in a.w, b, c), the parts of
  // a composite shape; SHAPE_TRIANGLE records index them by `off`
  const float4* __restrict__ tri;
};

// A part of a composite shape: triangle `off` of Pool::tri (TriMesh::triangle_at,
// src/shape/trimesh.rs:471-475), in the mesh's frame.
#define SHAPE_TRIANGLE 5u
// The core segment of a capsule (Capsule::segment, src/shape/capsule.rs:54-61): a = (0, -a, 0),
// b = (0, a, 0) with `a` the capsule's half height.  The contact-manifold generators replace a
// capsule by it (capsule_shape_manifold_generator.rs:34-46).
#define SHAPE_SEGMENT 6u

__device__ __forceinline__ void tri_vertices(const Pool& P, uint32_t k, f3& a, f3& b, f3& c) {
  const float4* t = P.tri + 3 * (size_t)k;
  const float4 A = t[0], B = t[1], C = t[2];
  a = f3{A.x, A.y, A.z};
  b = f3{B.x, B.y, B.z};
  c = f3{C.x, C.y, C.z};
}

// point_cloud_support_point_id (src/utils/point_cloud_support_point.rs:6-24), split over G lanes.
// The reference keeps index 0 when dot(p_0, dir) is NaN and otherwise returns the first index of
// the maximum over the non-NaN dots (NaN never compares greater).  Each lane scans vertices
// gl, gl + G, ... in increasing order keeping its first maximum over non-NaN dots; the lane
// reduction keeps the larger value and, on ties, the lower index.  U vertices' loads are issued
// together so their latencies overlap.
template <int G>
__device__ __forceinline__ uint32_t hull_support_id(const Pool& P, uint32_t off, uint32_t cnt, f3 dir, int gl) {
  constexpr int U = 4;
  constexpr uint32_t NONE = 0xFFFFFFFFu;
  float best = 0.f;
  uint32_t bi = NONE;
  int nan0 = 0;
  for (uint32_t v0 = gl; v0 < cnt; v0 += U * G) {
    float px[U], py[U], pz[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      uint32_t v = v0 + u * G;
      uint32_t vv = off + (v < cnt ? v : v0);  // always a vertex of this hull
      px[u] = P.x[vv];
      py[u] = P.y[vv];
      pz[u] = P.z[vv];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      uint32_t v = v0 + u * G;
      float d = dot(f3{px[u], py[u], pz[u]}, dir);
      if (v < cnt) {
        if (v == 0) nan0 = d != d;
        if (d > best || (bi == NONE && d == d)) {
          best = d;
          bi = v;
        }
      }
    }
  }
#pragma unroll
  for (int o = G / 2; o >= 1; o >>= 1) {
    float ob = __shfl_xor(best, o, 64);
    uint32_t oi = __shfl_xor(bi, o, 64);
    if (oi != NONE && (bi == NONE || ob > best || (ob == best && oi < bi))) {
      best = ob;
      bi = oi;
    }
  }
  nan0 = __shfl(nan0, 0, G);  // lane 0 of the group scanned vertex 0
  return (bi == NONE || nan0) ? 0u : bi;
}

// SupportMap::support_point (src/shape/support_map.rs:26-29) and the overrides of
// Ball (ball.rs:31-38), Cuboid (cuboid.rs:135-146), Capsule (capsule.rs:72-84),
// ConvexHull (convex.rs:418-423), ConstantOrigin (special_support_maps.rs:9-13).
template <int G>
__device__ __forceinline__ f3 support(const ShapeD& g, const iso& m, f3 dir, const Pool& P, int gl) {
  switch (g.type) {
    case NCX_BALL: {
      f3 u = normalize(dir);
      return m.t + u * g.a;
    }
    case NCX_CUBOID: {
      f3 ld = inv_xform_vec(m, dir);
      f3 p = f3{copysignf(g.a, ld.x), copysignf(g.b, ld.y), copysignf(g.c, ld.z)};
      return xform_pt(m, p);
    }
    case NCX_CAPSULE: {
      f3 ld = inv_xform_vec(m, dir);
      f3 u = normalize(ld);
      f3 res = f3{0.f, copysignf(g.a, u.y), 0.f};
      return xform_pt(m, res + u * g.b);
    }
    case NCX_CONVEX_HULL: {
      f3 ld = inv_xform_vec(m, dir);
      uint32_t id = hull_support_id<G>(P, g.off, g.cnt, ld, gl);
      f3 p = f3{P.x[g.off + id], P.y[g.off + id], P.z[g.off + id]};
      return xform_pt(m, p);
    }
    case SHAPE_TRIANGLE: {  // Triangle::local_support_point (src/shape/triangle.rs:290-307)
      f3 ld = inv_xform_vec(m, dir);
      f3 a, b, c;
      tri_vertices(P, g.off, a, b, c);
      const float d1 = dot(a, ld), d2 = dot(b, ld), d3 = dot(c, ld);
      const f3 p = d1 > d2 ? (d1 > d3 ? a : c) : (d2 > d3 ? b : c);
      return xform_pt(m, p);
    }
    case SHAPE_SEGMENT: {  // Segment::local_support_point (src/shape/segment.rs:182-191)
      f3 ld = inv_xform_vec(m, dir);
      const f3 a = f3{0.f, -g.a, 0.f}, b = f3{0.f, g.a, 0.f};
      return xform_pt(m, dot(a, ld) > dot(b, ld) ? a : b);
    }
    default:
      return f3{0.f, 0.f, 0.f};
  }
}

struct CSO {
  f3 p, o1, o2;
};

struct Ctx2 {
  iso m1, m2;
  ShapeD g1, g2;
};

template <int G>
__device__ __forceinline__ CSO cso_from(const Ctx2& c, f3 dir, const Pool& P, int gl) {
  CSO r;
  r.o1 = support<G>(c.g1, c.m1, dir, P, gl);
  r.o2 = support<G>(c.g2, c.m2, -dir, P, gl);
  r.p = r.o1 - r.o2;
  return r;
}

// ------------------------------------------------------------------ Voronoi projections of the origin
#define LOC_VERTEX 0
#define LOC_EDGE 1
#define LOC_FACE 2
#define LOC_SOLID 3
struct Loc {
  int kind, id;
  float bc0, bc1, bc2;
};

// point_segment.rs:52-91 (identity isometry, point = origin)
__device__ __forceinline__ f3 proj_segment(f3 a, f3 b, Loc& loc) {
  iso m = iso_identity();
  f3 ls = inv_xform_pt(m, f3{0.f, 0.f, 0.f});
  f3 ab = b - a, ap = ls - a;
  float ab_ap = dot(ab, ap), sqnab = norm_sq(ab);
  if (ab_ap <= 0.f) { loc.kind = LOC_VERTEX; loc.id = 0; return xform_pt(m, a); }
  if (ab_ap >= sqnab) { loc.kind = LOC_VERTEX; loc.id = 1; return xform_pt(m, b); }
  float u = ab_ap / sqnab;
  loc.kind = LOC_EDGE; loc.id = 0; loc.bc0 = 1.f - u; loc.bc1 = u;
  return xform_pt(m, a + ab * u);
}

// point_triangle.rs:61-309, dim3, solid = true
__device__ __forceinline__ f3 proj_triangle(f3 a, f3 b, f3 c, Loc& loc) {
  iso m = iso_identity();
  f3 o = f3{0.f, 0.f, 0.f};
  f3 p = inv_xform_pt(m, o);
  f3 ab = b - a, ac = c - a, ap = p - a;
  float ab_ap = dot(ab, ap), ac_ap = dot(ac, ap);
  if (ab_ap <= 0.f && ac_ap <= 0.f) { loc.kind = LOC_VERTEX; loc.id = 0; return xform_pt(m, a); }
  f3 bp = p - b;
  float ab_bp = dot(ab, bp), ac_bp = dot(ac, bp);
  if (ab_bp >= 0.f && ac_bp <= ab_bp) { loc.kind = LOC_VERTEX; loc.id = 1; return xform_pt(m, b); }
  f3 cp = p - c;
  float ab_cp = dot(ab, cp), ac_cp = dot(ac, cp);
  if (ac_cp >= 0.f && ab_cp <= ac_cp) { loc.kind = LOC_VERTEX; loc.id = 2; return xform_pt(m, c); }
  f3 bc = c - b;
  f3 n = cross(ab, ac);
  float vc = dot(n, cross(ab, ap));
  if (vc < 0.f && ab_ap >= 0.f && ab_bp <= 0.f) {
    float v = ab_ap / norm_sq(ab);
    loc.kind = LOC_EDGE; loc.id = 0; loc.bc0 = 1.f - v; loc.bc1 = v;
    return xform_pt(m, a + ab * v);
  }
  float vb = -dot(n, cross(ac, cp));
  if (vb < 0.f && ac_ap >= 0.f && ac_cp <= 0.f) {
    float w = ac_ap / norm_sq(ac);
    loc.kind = LOC_EDGE; loc.id = 2; loc.bc0 = 1.f - w; loc.bc1 = w;
    return xform_pt(m, a + ac * w);
  }
  float va = dot(n, cross(bc, bp));
  if (va < 0.f && ac_bp - ab_bp >= 0.f && ab_cp - ac_cp >= 0.f) {
    float w = dot(bc, bp) / norm_sq(bc);
    loc.kind = LOC_EDGE; loc.id = 1; loc.bc0 = 1.f - w; loc.bc1 = w;
    return xform_pt(m, b + bc * w);
  }
  int clockwise = dot(n, ap) >= 0.f ? 0 : 1;
  if (va + vb + vc != 0.f) {
    float denom = 1.f / (va + vb + vc);
    float v = vb * denom, w = vc * denom;
    loc.kind = LOC_FACE; loc.id = clockwise; loc.bc0 = 1.f - v - w; loc.bc1 = v; loc.bc2 = w;
    return xform_pt(m, a + ab * v + ac * w);
  }
  loc.kind = LOC_SOLID; loc.id = 0;
  return o;
}

// check_edge of point_tetrahedron.rs:106-152
__device__ __forceinline__ bool tet_edge(int i, const iso& m, f3 a, f3 nabc, f3 nabd, f3 ap, f3 ab, float ap_ab,
                                         float bp_ab, float& dabc, float& dabd, Loc& loc, f3& out) {
  float ab_ab = ap_ab - bp_ab;
  f3 ap_x_ab = cross(ap, ab);
  dabc = dot(ap_x_ab, nabc);
  dabd = dot(ap_x_ab, nabd);
  if (ab_ab != 0.f && dabc >= 0.f && dabd >= 0.f && ap_ab >= 0.f && ap_ab <= ab_ab) {
    float u = ap_ab / ab_ab;
    loc.kind = LOC_EDGE; loc.id = i; loc.bc0 = 1.f - u; loc.bc1 = u;
    out = xform_pt(m, a + ab * u);
    return true;
  }
  return false;
}

// check_face of point_tetrahedron.rs:251-306
__device__ __forceinline__ bool tet_face(int i, const iso& m, f3 a, f3 b, f3 c, f3 ap, f3 bp, f3 cp, f3 ab, f3 ac,
                                         f3 ad, float d1, float d2, float d3, Loc& loc, f3& out) {
  if (d1 < 0.f && d2 < 0.f && d3 < 0.f) {
    f3 n = cross(ab, ac);
    if (dot(n, ad) * dot(n, ap) < 0.f) {
      f3 nn;
      float nl;
      if (!try_normalize(n, NCX_FLT_EPS, &nn, &nl)) return false;
      float vc = dot(nn, cross(ap, bp));
      float va = dot(nn, cross(bp, cp));
      float vb = dot(nn, cross(cp, ap));
      float denom = va + vb + vc;
      float inv_denom = 1.f / denom;
      float b0 = va * inv_denom, b1 = vb * inv_denom, b2 = vc * inv_denom;
      loc.kind = LOC_FACE; loc.id = i; loc.bc0 = b0; loc.bc1 = b1; loc.bc2 = b2;
      out = xform_pt(m, a * b0 + b * b1 + c * b2);
      return true;
    }
  }
  return false;
}

// point_tetrahedron.rs:35-353, solid = true
__device__ __forceinline__ f3 proj_tetrahedron(f3 A, f3 B, f3 C, f3 D, Loc& loc) {
  iso m = iso_identity();
  f3 p = inv_xform_pt(m, f3{0.f, 0.f, 0.f});
  f3 ab = B - A, ac = C - A, ad = D - A, ap = p - A;
  float ap_ab = dot(ap, ab), ap_ac = dot(ap, ac), ap_ad = dot(ap, ad);
  if (ap_ab <= 0.f && ap_ac <= 0.f && ap_ad <= 0.f) { loc.kind = LOC_VERTEX; loc.id = 0; return xform_pt(m, A); }
  f3 bc = C - B, bd = D - B, bp = p - B;
  float bp_bc = dot(bp, bc), bp_bd = dot(bp, bd), bp_ab = dot(bp, ab);
  if (bp_bc <= 0.f && bp_bd <= 0.f && bp_ab >= 0.f) { loc.kind = LOC_VERTEX; loc.id = 1; return xform_pt(m, B); }
  f3 cd = D - C, cp = p - C;
  float cp_ac = dot(cp, ac), cp_bc = dot(cp, bc), cp_cd = dot(cp, cd);
  if (cp_cd <= 0.f && cp_bc >= 0.f && cp_ac >= 0.f) { loc.kind = LOC_VERTEX; loc.id = 2; return xform_pt(m, C); }
  f3 dp = p - D;
  float dp_cd = dot(dp, cd), dp_bd = dot(dp, bd), dp_ad = dot(dp, ad);
  if (dp_ad >= 0.f && dp_bd >= 0.f && dp_cd >= 0.f) { loc.kind = LOC_VERTEX; loc.id = 3; return xform_pt(m, D); }
  f3 out;
  float dabc, dabd, dacd, dacb, dadb, dadc, dbca, dbcd, dbdc, dbda, dcda, dcdb;
  f3 nabc = cross(ab, ac), nabd = cross(ab, ad);
  if (tet_edge(0, m, A, nabc, nabd, ap, ab, ap_ab, bp_ab, dabc, dabd, loc, out)) return out;
  f3 nacd = cross(ac, ad);
  if (tet_edge(1, m, A, nacd, -nabc, ap, ac, ap_ac, cp_ac, dacd, dacb, loc, out)) return out;
  if (tet_edge(2, m, A, -nabd, -nacd, ap, ad, ap_ad, dp_ad, dadb, dadc, loc, out)) return out;
  f3 nbcd = cross(bc, bd);
  if (tet_edge(3, m, B, nabc, nbcd, bp, bc, bp_bc, cp_bc, dbca, dbcd, loc, out)) return out;
  if (tet_edge(4, m, B, -nbcd, nabd, bp, bd, bp_bd, dp_bd, dbdc, dbda, loc, out)) return out;
  if (tet_edge(5, m, C, nacd, nbcd, cp, cd, cp_cd, dp_cd, dcda, dcdb, loc, out)) return out;
  if (tet_face(0, m, A, B, C, ap, bp, cp, ab, ac, ad, dabc, dbca, dacb, loc, out)) return out;
  if (tet_face(1, m, A, B, D, ap, bp, dp, ab, ad, ac, dadb, dabd, dbda, loc, out)) return out;
  if (tet_face(2, m, A, C, D, ap, cp, dp, ac, ad, ab, dacd, dcda, dadc, loc, out)) return out;
  if (tet_face(3, m, B, C, D, bp, cp, dp, bc, bd, -ab, dbcd, dcdb, dbdc, loc, out)) return out;
  loc.kind = LOC_SOLID; loc.id = 0;
  return xform_pt(m, p);
}

// ------------------------------------------------------------------ VoronoiSimplex (voronoi_simplex3.rs:12-352)
// The four vertices are separate members and the "previous vertex" permutation is
// packed 2 bits per slot, so no runtime index reaches an array: everything stays in
// registers.
struct Simplex {
  CSO v0, v1, v2, v3;
  uint32_t pv;  // prev_vertices, slot i in bits [2i, 2i+2)
  float proj0, proj1, proj2, pproj0, pproj1, pproj2;
  int dim, pdim;
};

// Value selects (never a select between addresses), so the simplex stays in registers.
__device__ __forceinline__ f3 sel3(bool c, f3 a, f3 b) { return f3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }
__device__ __forceinline__ CSO csel(bool c, const CSO& a, const CSO& b) {
  return CSO{sel3(c, a.p, b.p), sel3(c, a.o1, b.o1), sel3(c, a.o2, b.o2)};
}
__device__ __forceinline__ CSO sx_get(const Simplex& s, int i) {
  return csel(i == 0, s.v0, csel(i == 1, s.v1, csel(i == 2, s.v2, s.v3)));
}
__device__ __forceinline__ void sx_set(Simplex& s, int i, const CSO& c) {
  s.v0 = csel(i == 0, c, s.v0);
  s.v1 = csel(i == 1, c, s.v1);
  s.v2 = csel(i == 2, c, s.v2);
  s.v3 = csel(i == 3, c, s.v3);
}
__device__ __forceinline__ void sx_swap(Simplex& s, int i, int j) {
  CSO a = sx_get(s, i), b = sx_get(s, j);
  sx_set(s, i, b);
  sx_set(s, j, a);
  uint32_t pi = (s.pv >> (2 * i)) & 3u, pj = (s.pv >> (2 * j)) & 3u;
  s.pv &= ~((3u << (2 * i)) | (3u << (2 * j)));
  s.pv |= (pj << (2 * i)) | (pi << (2 * j));
}

#define PV_IDENTITY 0xE4u  // slots (0, 1, 2, 3)

__device__ __forceinline__ void sx_init(Simplex& s) {
  s.pv = PV_IDENTITY;
  s.proj0 = s.proj1 = s.proj2 = 0.f;
  s.pproj0 = s.pproj1 = s.pproj2 = 0.f;
  s.dim = 0;
  s.pdim = 0;
}

__device__ __forceinline__ void sx_reset(Simplex& s, const CSO& p) {
  s.dim = 0;
  s.pdim = 0;
  s.v0 = p;
}

__device__ __forceinline__ bool sx_add(Simplex& s, const CSO& pt) {
  const float tol = NCX_FLT_EPS * 10.f;
  s.pdim = s.dim;
  s.pproj0 = s.proj0; s.pproj1 = s.proj1; s.pproj2 = s.proj2;
  s.pv = PV_IDENTITY;
  if (s.dim == 0) {
    if (norm_sq(s.v0.p - pt.p) < tol) return false;
  } else if (s.dim == 1) {
    f3 ab = s.v1.p - s.v0.p, ac = pt.p - s.v0.p;
    if (norm_sq(cross(ab, ac)) < tol) return false;
  } else {
    f3 ab = s.v1.p - s.v0.p, ac = s.v2.p - s.v0.p, ap = pt.p - s.v0.p;
    f3 n = normalize(cross(ab, ac));
    if (fabsf(dot(n, ap)) < tol) return false;
  }
  s.dim += 1;
  sx_set(s, s.dim, pt);
  return true;
}

__device__ __forceinline__ f3 sx_project_reduce(Simplex& s) {
  if (s.dim == 0) {
    s.proj0 = 1.f;
    return s.v0.p;
  } else if (s.dim == 1) {
    Loc loc;
    f3 pr = proj_segment(s.v0.p, s.v1.p, loc);
    if (loc.kind == LOC_VERTEX && loc.id == 0) { s.proj0 = 1.f; s.dim = 0; }
    else if (loc.kind == LOC_VERTEX) { sx_swap(s, 0, 1); s.proj0 = 1.f; s.dim = 0; }
    else { s.proj0 = loc.bc0; s.proj1 = loc.bc1; }
    return pr;
  } else if (s.dim == 2) {
    Loc loc;
    f3 pr = proj_triangle(s.v0.p, s.v1.p, s.v2.p, loc);
    if (loc.kind == LOC_VERTEX) { sx_swap(s, 0, loc.id); s.proj0 = 1.f; s.dim = 0; }
    else if (loc.kind == LOC_EDGE) {
      if (loc.id == 0) { s.proj0 = loc.bc0; s.proj1 = loc.bc1; }
      else if (loc.id == 1) { sx_swap(s, 0, 2); s.proj0 = loc.bc1; s.proj1 = loc.bc0; }
      else { sx_swap(s, 1, 2); s.proj0 = loc.bc0; s.proj1 = loc.bc1; }
      s.dim = 1;
    } else if (loc.kind == LOC_FACE) {
      s.proj0 = loc.bc0; s.proj1 = loc.bc1; s.proj2 = loc.bc2;
    }
    return pr;
  } else {
    Loc loc;
    f3 pr = proj_tetrahedron(s.v0.p, s.v1.p, s.v2.p, s.v3.p, loc);
    if (loc.kind == LOC_VERTEX) { sx_swap(s, 0, loc.id); s.proj0 = 1.f; s.dim = 0; }
    else if (loc.kind == LOC_EDGE) {
      switch (loc.id) {
        case 1: sx_swap(s, 1, 2); break;
        case 2: sx_swap(s, 1, 3); break;
        case 3: sx_swap(s, 0, 2); break;
        case 4: sx_swap(s, 0, 3); break;
        case 5: sx_swap(s, 0, 2); sx_swap(s, 1, 3); break;
        default: break;
      }
      if (loc.id == 3 || loc.id == 4) { s.proj0 = loc.bc1; s.proj1 = loc.bc0; }
      else { s.proj0 = loc.bc0; s.proj1 = loc.bc1; }
      s.dim = 1;
    } else if (loc.kind == LOC_FACE) {
      const int fid = loc.id;
      s.v2 = csel(fid == 1, s.v3, s.v2);
      s.v1 = csel(fid == 2, s.v3, s.v1);
      s.v0 = csel(fid >= 3, s.v3, s.v0);
      switch (fid) {
        case 0: s.proj0 = loc.bc0; s.proj1 = loc.bc1; s.proj2 = loc.bc2; break;
        case 1: s.proj0 = loc.bc0; s.proj1 = loc.bc1; s.proj2 = loc.bc2; break;
        case 2: s.proj0 = loc.bc0; s.proj1 = loc.bc2; s.proj2 = loc.bc1; break;
        default: s.proj0 = loc.bc2; s.proj1 = loc.bc0; s.proj2 = loc.bc1; break;
      }
      s.dim = 2;
    }
    return pr;
  }
}

// gjk.rs:367-388
__device__ __forceinline__ void gjk_result(const Simplex& s, bool prev, f3& p1, f3& p2) {
  f3 r1 = f3{0.f, 0.f, 0.f}, r2 = f3{0.f, 0.f, 0.f};
  const int n = prev ? s.pdim : s.dim;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i <= n) {
      float c = prev ? (i == 0 ? s.pproj0 : (i == 1 ? s.pproj1 : s.pproj2))
                     : (i == 0 ? s.proj0 : (i == 1 ? s.proj1 : s.proj2));
      CSO pt = sx_get(s, prev ? (int)((s.pv >> (2 * i)) & 3u) : i);
      r1 = r1 + pt.o1 * c;
      r2 = r2 + pt.o2 * c;
    }
  }
  p1 = r1;
  p2 = r2;
}

struct GJKOut {
  int status;
  f3 p1, p2, dir;
  int iters;
};

// gjk::closest_points (gjk.rs:76-177), exact_dist = true.
template <int G>
__device__ __forceinline__ GJKOut gjk(const Ctx2& c, float max_dist, Simplex& s, const Pool& P, int gl, int max_iter) {
  GJKOut r;
  r.iters = 0;
  const float eps_tol = NCX_FLT_EPS * 10.f;
  const float eps_rel = __builtin_sqrtf(eps_tol);
  f3 proj = sx_project_reduce(s);
  f3 old_dir;
  {
    f3 pd;
    float nl;
    if (try_normalize(proj, 0.f, &pd, &nl)) old_dir = -pd;
    else { r.status = GJK_S_INTERSECTION; return r; }
  }
  float max_bound = NCX_FLT_MAX;
  f3 dir;
  int niter = 0;
  for (;;) {
    float old_max_bound = max_bound;
    {
      f3 nd;
      float dist;
      if (try_normalize(-proj, eps_tol, &nd, &dist)) { dir = nd; max_bound = dist; }
      else { r.status = GJK_S_INTERSECTION; return r; }
    }
    if (max_bound >= old_max_bound) {
      gjk_result(s, true, r.p1, r.p2);
      r.dir = old_dir;
      r.status = GJK_S_CLOSEST;
      return r;
    }
    CSO cso = cso_from<G>(c, dir, P, gl);
    r.iters += 1;
    float min_bound = -dot(dir, cso.p);
    if (!(min_bound == min_bound)) { r.status = GJK_S_NAN; return r; }
    if (min_bound > max_dist) { r.dir = dir; r.status = GJK_S_NOINTER; return r; }
    else if (max_bound - min_bound <= eps_rel * max_bound) {
      gjk_result(s, false, r.p1, r.p2);
      r.dir = dir;
      r.status = GJK_S_CLOSEST;
      return r;
    }
    if (!sx_add(s, cso)) {
      gjk_result(s, false, r.p1, r.p2);
      r.dir = dir;
      r.status = GJK_S_CLOSEST;
      return r;
    }
    old_dir = dir;
    proj = sx_project_reduce(s);
    if (s.dim == 3) {
      if (min_bound >= eps_tol) {
        gjk_result(s, true, r.p1, r.p2);
        r.dir = old_dir;
        r.status = GJK_S_CLOSEST;
        return r;
      }
      r.status = GJK_S_INTERSECTION;
      return r;
    }
    niter += 1;
    if (niter == max_iter) { r.dir = f3{1.f, 0.f, 0.f}; r.status = GJK_S_CAP; return r; }
  }
}

// ------------------------------------------------------------------ time of impact and ray casts on support maps
// approx::relative_eq!(x, 0) with the default epsilon and max_relative: |x| <= f32::EPSILON.
__device__ __forceinline__ bool relative_eq_zero(float x) { return fabsf(x) <= NCX_FLT_EPS; }

// SupportMap::support_point_toward (src/shape/support_map.rs:31-34): `u` is unit and used as given.
// Ball overrides it (ball.rs:36-39); Capsule's local_support_point_toward does not normalise
// (capsule.rs:79-83); Cuboid and ConvexHull fall back to support_point.
template <int G>
__device__ __forceinline__ f3 support_toward(const ShapeD& g, const iso& m, f3 u, const Pool& P, int gl) {
  if (g.type == NCX_BALL) return m.t + u * g.a;
  if (g.type == NCX_CAPSULE) {
    f3 ld = inv_xform_vec(m, u);
    f3 res = f3{0.f, copysignf(g.a, ld.y), 0.f};
    return xform_pt(m, res + ld * g.b);
  }
  return support<G>(g, m, u, P, gl);
}

// Two support maps, the first dilated by `dil` when `dilated` (DilatedShape,
// src/query/algorithms/special_support_maps.rs:32-59: support_point normalises the direction and
// adds the ball's support point to the shape's).
struct RCtx {
  iso m1, m2;
  ShapeD g1, g2;
  float dil;
  bool dilated;
};

template <int G>
__device__ __forceinline__ CSO cso_from_r(const RCtx& c, f3 dir, const Pool& P, int gl) {
  CSO r;
  if (c.dilated) {
    f3 u = normalize(dir);
    r.o1 = support_toward<G>(c.g1, c.m1, u, P, gl) + u * c.dil;
  } else {
    r.o1 = support<G>(c.g1, c.m1, dir, P, gl);
  }
  r.o2 = support<G>(c.g2, c.m2, -dir, P, gl);
  r.p = r.o1 - r.o2;
  return r;
}

__device__ __forceinline__ bool same_bits(float a, float b) { return __float_as_uint(a) == __float_as_uint(b); }
__device__ __forceinline__ bool same_bits3(f3 a, f3 b) {
  return same_bits(a.x, b.x) && same_bits(a.y, b.y) && same_bits(a.z, b.z);
}

// gjk::minkowski_ray_cast (src/query/algorithms/gjk.rs:245-365) on m1*g1 - m2*g2, ray (o, d).
// Returns false for None.  The improved_fixed_point_support feature is off (the crate's default).
// An iteration that leaves every value the loop reads bit-identical (the simplex's points and
// dimension, the ray origin and toi, the projection, max_bound, last_chance and the last normal)
// is repeated identically until the 10000-iteration cap, so it returns None at once: a shift too
// small to change the origin (its f32 sums absorb it) with a support point the simplex rejects
// as affinely dependent is such a fixed point.
template <int G>
__device__ __forceinline__ bool minkowski_ray_cast(const RCtx& c, f3 ro, f3 rd, float max_toi, Simplex& s,
                                                   const Pool& P, int gl, float& toi_out, f3& normal_out) {
  const float eps_tol = NCX_FLT_EPS * 10.f;
  const float eps_rel = __builtin_sqrtf(eps_tol);
  const float ray_length = norm(rd);
  if (relative_eq_zero(ray_length)) return false;
  float ltoi = 0.f;
  f3 co = ro, cd = rd / ray_length;
  f3 dir = -cd;
  f3 ldir = dir;
  CSO sp = cso_from_r<G>(c, dir, P, gl);
  {
    CSO t0 = sp;
    t0.p = sp.p + (-co);
    sx_reset(s, t0);
  }
  f3 proj = sx_project_reduce(s);
  float max_bound = NCX_FLT_MAX;
  int niter = 0;
  bool last_chance = false;
  for (;;) {
    const float old_max_bound = max_bound;
    const f3 proj_in = proj;
    const bool lc_in = last_chance;
    const int dim_in = s.dim;
    bool same = true;  // nothing changed so far in this iteration
    {
      f3 nd;
      float dist;
      if (try_normalize(-proj, eps_tol, &nd, &dist)) { dir = nd; max_bound = dist; }
      else { toi_out = ltoi / ray_length; normal_out = ldir; return true; }
    }
    if (max_bound >= old_max_bound) {
      last_chance = true;  // CSOPoint::single_point (cso_point.rs:42-44)
      f3 p = proj + co;
      sp.p = p; sp.o1 = p; sp.o2 = f3{0.f, 0.f, 0.f};
    } else {
      sp = cso_from_r<G>(c, dir, P, gl);
    }
    if (last_chance && ltoi > 0.f) { toi_out = ltoi / ray_length; normal_out = ldir; return true; }
    // query::ray_toi_with_plane(support point, dir, curr_ray) (src/query/ray/ray_plane.rs:8-37)
    const float denom = dot(dir, cd);
    bool has_t = false;
    float t = 0.f;
    if (!relative_eq_zero(denom)) {
      t = dot(dir, sp.p - co) / denom;
      has_t = t >= 0.f;
    }
    if (has_t) {
      if (denom < 0.f && t > 0.f) {
        same = same_bits3(ldir, dir);
        ldir = dir;
        const float ltoi_in = ltoi;
        ltoi += t;
        if (ltoi / ray_length > max_toi) return false;
        const f3 shift = cd * t;
        const f3 co_in = co;
        co = co + shift;
        max_bound = NCX_FLT_MAX;
        const f3 ns = -shift;  // VoronoiSimplex::modify_pnts over vertices 0..=dim (voronoi_simplex3.rs:347-351)
        const f3 v0 = s.v0.p + ns, v1 = s.v1.p + ns, v2 = s.v2.p + ns, v3 = s.v3.p + ns;
        same = same && same_bits(ltoi, ltoi_in) && same_bits3(co, co_in) && same_bits3(v0, s.v0.p) &&
               (s.dim < 1 || same_bits3(v1, s.v1.p)) && (s.dim < 2 || same_bits3(v2, s.v2.p)) &&
               (s.dim < 3 || same_bits3(v3, s.v3.p));
        s.v0.p = v0;
        s.v1.p = sel3(s.dim >= 1, v1, s.v1.p);
        s.v2.p = sel3(s.dim >= 2, v2, s.v2.p);
        s.v3.p = sel3(s.dim >= 3, v3, s.v3.p);
        last_chance = false;
      }
    } else if (denom > eps_tol) {
      return false;
    }
    if (last_chance) return false;
    const float min_bound = -dot(dir, sp.p - co);
    if (max_bound - min_bound <= eps_rel * max_bound) return false;
    bool added;
    {
      CSO tp = sp;
      tp.p = sp.p + (-co);
      added = sx_add(s, tp);
    }
    proj = sx_project_reduce(s);
    if (s.dim == 3) {
      if (min_bound >= eps_tol) return false;
      toi_out = ltoi / ray_length;
      normal_out = ldir;
      return true;
    }
    // a rejected point leaves the points in place, and a reduction that keeps the dimension of
    // a point, segment or triangle keeps them in order
    if (same && !added && s.dim == dim_in && last_chance == lc_in && same_bits(max_bound, old_max_bound) &&
        same_bits3(proj, proj_in))
      return false;
    niter += 1;
    if (niter == 10000) return false;
  }
}

// ray_intersection_with_support_map_with_params (src/query/ray/ray_support_map.rs:15-62), shape
// at the identity (the callers pass a local-frame ray).
template <int G>
__device__ __forceinline__ bool ray_support_map(const ShapeD& g, f3 ro, f3 rd, float max_toi, bool solid,
                                                const Pool& P, int gl, float& toi, f3& normal) {
  RCtx c;
  c.m1 = iso_identity();
  c.m2 = iso_identity();
  c.g1 = g;
  c.g2.type = SHAPE_ORIGIN;
  c.dil = 0.f;
  c.dilated = false;
  Simplex s;
  sx_init(s);
  float t;
  f3 n;
  if (!minkowski_ray_cast<G>(c, ro, rd, max_toi, s, P, gl, t, n)) return false;
  if (solid || t != 0.f) { toi = t; normal = n; return true; }
  // the ray starts inside the shape: cast back from a point past its far side
  const f3 ndir = normalize(rd);
  const f3 supp = support<G>(g, c.m1, ndir, P, gl);
  const float eps = 0.001f;
  const float shift = dot(supp - ro, ndir) + eps;
  Simplex s2;
  sx_init(s2);
  if (!minkowski_ray_cast<G>(c, ro + ndir * shift, -rd, shift + eps, s2, P, gl, t, n)) return false;
  const float tt = shift - t;
  if (!(tt <= max_toi)) return false;
  toi = tt;
  normal = n;
  return true;
}

// ray_toi_with_ball (src/query/ray/ray_ball.rs:83-124).
__device__ __forceinline__ bool ray_toi_with_ball(f3 center, float radius, f3 ro, f3 rd, bool solid, bool& inside,
                                                  float& toi) {
  const f3 dc = ro - center;
  const float a = norm_sq(rd);
  const float b = dot(dc, rd);
  const float c = norm_sq(dc) - radius * radius;
  inside = false;
  if (a == 0.f) {
    if (c > 0.f) return false;
    inside = true;
    toi = 0.f;
    return true;
  }
  if (c > 0.f && b > 0.f) return false;
  const float delta = b * b - a * c;
  if (delta < 0.f) return false;
  const float t = (-b - __builtin_sqrtf(delta)) / a;
  if (t <= 0.f) {
    inside = true;
    toi = solid ? 0.f : (-b + __builtin_sqrtf(delta)) / a;
    return true;
  }
  toi = t;
  return true;
}

#define NCX_FEAT_UNKNOWN 0xFFFFFFFEu

// AABB ray cast with normals: clip_line (src/query/ray/ray_aabb.rs:178-268) and ray_aabb
// (:270-290), local frame; side is +(i+1) / -(i+1) as in the reference.
__device__ __forceinline__ bool ray_box_normal(f3 h, f3 ro, f3 rd, float max_toi, bool solid, float& toi, f3& normal,
                                               int& side) {
  float tmax = NCX_FLT_MAX;
  float tmin = -tmax;
  int near_side = 0, far_side = 0;
  bool near_diag = false, far_diag = false;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float di = comp(rd, i), oi = comp(ro, i), hi = comp(h, i);
    if (di == 0.f) {
      if (oi < -hi || oi > hi) return false;
    } else {
      const float denom = 1.f / di;
      float tn = (-hi - oi) * denom;
      float tf = (hi - oi) * denom;
      const bool flip = tn > tf;
      if (flip) { float x = tn; tn = tf; tf = x; }
      if (tn > tmin) { tmin = tn; near_side = flip ? -(i + 1) : i + 1; near_diag = false; }
      else if (tn == tmin) near_diag = true;
      if (tf < tmax) { tmax = tf; far_side = !flip ? -(i + 1) : i + 1; far_diag = false; }
      else if (tf == tmax) far_diag = true;
      if (tmax < 0.f || tmin > tmax) return false;
    }
  }
  f3 nn = f3{0.f, 0.f, 0.f}, nf = nn;
  if (near_diag) nn = -normalize(rd);
  else if (near_side < 0) set_comp(nn, -near_side - 1, 1.f);
  else set_comp(nn, near_side - 1, -1.f);
  if (far_diag) nf = -normalize(rd);
  else if (far_side < 0) set_comp(nf, -far_side - 1, -1.f);
  else set_comp(nf, far_side - 1, 1.f);
  if (tmin < 0.f) {
    if (solid) { toi = 0.f; normal = f3{0.f, 0.f, 0.f}; side = far_side; return true; }
    if (tmax <= max_toi) { toi = tmax; normal = nf; side = far_side; return true; }
    return false;
  }
  if (tmin <= max_toi) { toi = tmin; normal = nn; side = near_side; return true; }
  return false;
}

// RayCast::toi_and_normal_with_ray of Ball (src/query/ray/ray_ball.rs:32-43, 127-142), Cuboid
// (ray_cuboid.rs:17-26 -> ray_aabb.rs:45-61), Capsule and ConvexHull (ray_support_map.rs:117-163:
// local-frame ray, normal rotated back).  feature: face id, 0 for a ball, NCX_FEAT_UNKNOWN for
// support-map shapes.
// The closed forms (Ball, Cuboid) of shape_ray_cast below; no cross-lane work, so one lane can
// cast a ray alone.  A shape of another type gives no hit.
__device__ __forceinline__ bool shape_ray_cast_closed(const ShapeD& g, const iso& m, f3 ro, f3 rd, float max_toi,
                                                      bool solid, float& toi, f3& normal, uint32_t& feature) {
  if (g.type == NCX_BALL) {
    bool inside;
    float t;
    if (!ray_toi_with_ball(m.t, g.a, ro, rd, solid, inside, t)) return false;
    const f3 n = normalize((ro + rd * t) - m.t);
    if (!(t <= max_toi)) return false;
    toi = t;
    normal = inside ? -n : n;
    feature = 0;
    return true;
  }
  if (g.type == NCX_CUBOID) {
    const f3 lo = inv_xform_pt(m, ro), ld = inv_xform_vec(m, rd);
    float t;
    f3 n;
    int side;
    if (!ray_box_normal(f3{g.a, g.b, g.c}, lo, ld, max_toi, solid, t, n, side)) return false;
    toi = t;
    normal = xform_vec(m, n);
    feature = side < 0 ? (uint32_t)(-side - 1 + 3) : (uint32_t)(side - 1);
    return true;
  }
  return false;
}

template <int G>
__device__ __forceinline__ bool shape_ray_cast(const ShapeD& g, const iso& m, f3 ro, f3 rd, float max_toi, bool solid,
                                               const Pool& P, int gl, float& toi, f3& normal, uint32_t& feature) {
  if (g.type == NCX_BALL || g.type == NCX_CUBOID)
    return shape_ray_cast_closed(g, m, ro, rd, max_toi, solid, toi, normal, feature);
  if (g.type == NCX_CAPSULE || g.type == NCX_CONVEX_HULL) {
    const f3 lo = inv_xform_pt(m, ro), ld = inv_xform_vec(m, rd);
    float t;
    f3 n;
    if (!ray_support_map<G>(g, lo, ld, max_toi, solid, P, gl, t, n)) return false;
    toi = t;
    normal = xform_vec(m, n);
    feature = NCX_FEAT_UNKNOWN;
    return true;
  }
  return false;
}

// ------------------------------------------------------------------ triangles and isometries for composite shapes
// approx::relative_eq!(x, y) with the default epsilon and max_relative (f32::EPSILON).
__device__ __forceinline__ bool relative_eq_f(float x, float y) {
  if (x == y) return true;
  if (__builtin_isinf(x) || __builtin_isinf(y)) return false;
  const float d = fabsf(x - y);
  if (d <= NCX_FLT_EPS) return true;
  const float ax = fabsf(x), ay = fabsf(y);
  return d <= (ay > ax ? ay : ax) * NCX_FLT_EPS;
}

// Triangle::project_point_with_location with solid = true (src/query/point/point_triangle.rs:61-240)
// for the triangle (a, b, c) placed by m: the projection in the world frame; *inside is the
// PointProjection's is_inside (relative_eq!(projection, pt) in 3D, :6-19; true on the solid
// branch).
__device__ __forceinline__ f3 proj_triangle_at(const iso& m, f3 a, f3 b, f3 c, f3 pt, Loc& loc, bool& inside) {
  const f3 p = inv_xform_pt(m, pt);
  const f3 ab = b - a, ac = c - a, ap = p - a;
  const float ab_ap = dot(ab, ap), ac_ap = dot(ac, ap);
  f3 res;
  if (ab_ap <= 0.f && ac_ap <= 0.f) { loc.kind = LOC_VERTEX; loc.id = 0; res = xform_pt(m, a); }
  else {
    const f3 bp = p - b;
    const float ab_bp = dot(ab, bp), ac_bp = dot(ac, bp);
    if (ab_bp >= 0.f && ac_bp <= ab_bp) { loc.kind = LOC_VERTEX; loc.id = 1; res = xform_pt(m, b); }
    else {
      const f3 cp = p - c;
      const float ab_cp = dot(ab, cp), ac_cp = dot(ac, cp);
      if (ac_cp >= 0.f && ab_cp <= ac_cp) { loc.kind = LOC_VERTEX; loc.id = 2; res = xform_pt(m, c); }
      else {
        const f3 bc = c - b;
        const f3 n = cross(ab, ac);
        const float vc = dot(n, cross(ab, ap));
        const float vb = -dot(n, cross(ac, cp));
        const float va = dot(n, cross(bc, bp));
        if (vc < 0.f && ab_ap >= 0.f && ab_bp <= 0.f) {
          const float v = ab_ap / norm_sq(ab);
          loc.kind = LOC_EDGE; loc.id = 0; loc.bc0 = 1.f - v; loc.bc1 = v;
          res = xform_pt(m, a + ab * v);
        } else if (vb < 0.f && ac_ap >= 0.f && ac_cp <= 0.f) {
          const float w = ac_ap / norm_sq(ac);
          loc.kind = LOC_EDGE; loc.id = 2; loc.bc0 = 1.f - w; loc.bc1 = w;
          res = xform_pt(m, a + ac * w);
        } else if (va < 0.f && ac_bp - ab_bp >= 0.f && ab_cp - ac_cp >= 0.f) {
          const float w = dot(bc, bp) / norm_sq(bc);
          loc.kind = LOC_EDGE; loc.id = 1; loc.bc0 = 1.f - w; loc.bc1 = w;
          res = xform_pt(m, b + bc * w);
        } else if (va + vb + vc != 0.f) {
          const float denom = 1.f / (va + vb + vc);
          const float v = vb * denom, w = vc * denom;
          loc.kind = LOC_FACE; loc.id = dot(n, ap) >= 0.f ? 0 : 1;
          loc.bc0 = 1.f - v - w; loc.bc1 = v; loc.bc2 = w;
          res = xform_pt(m, a + ab * v + ac * w);
        } else {
          loc.kind = LOC_SOLID; loc.id = 0;
          inside = true;
          return pt;
        }
      }
    }
  }
  inside = relative_eq_f(res.x, pt.x) && relative_eq_f(res.y, pt.y) && relative_eq_f(res.z, pt.z);
  return res;
}

// Segment::project_point_with_location (src/query/point/point_segment.rs:49-91) for the segment
// (a, b) placed by m, and project_point_with_feature's feature (:13-41): vertex 0 or 1 (kind
// LOC_VERTEX) or the edge (LOC_EDGE).  *inside = relative_eq!(projection, pt).
__device__ __forceinline__ f3 proj_segment_at(const iso& m, f3 a, f3 b, f3 pt, int& kind, int& id, bool& inside) {
  const f3 ls = inv_xform_pt(m, pt);
  const f3 ab = b - a, ap = ls - a;
  const float ab_ap = dot(ab, ap), sqnab = norm_sq(ab);
  f3 proj;
  if (ab_ap <= 0.f) { kind = LOC_VERTEX; id = 0; proj = xform_pt(m, a); }
  else if (ab_ap >= sqnab) { kind = LOC_VERTEX; id = 1; proj = xform_pt(m, b); }
  else {
    const float u = ab_ap / sqnab;
    kind = LOC_EDGE; id = 0;
    proj = xform_pt(m, a + ab * u);
  }
  inside = relative_eq_f(proj.x, pt.x) && relative_eq_f(proj.y, pt.y) && relative_eq_f(proj.z, pt.z);
  return proj;
}

// Segment::feature_normal (src/shape/segment.rs:238-273), local frame, for the segment (a, b):
// vertex 0 -> its direction, vertex 1 -> minus it, the edge -> the unit vector along the
// direction's smallest component with the direction's part removed; +y when degenerate.
__device__ __forceinline__ f3 segment_feature_normal(f3 a, f3 b, int kind, int id) {
  f3 d;
  float nl;
  if (!try_normalize(b - a, NCX_FLT_EPS, &d, &nl)) return f3{0.f, 1.f, 0.f};
  if (kind == LOC_VERTEX) return id == 0 ? d : -d;
  // Vector::iamin: the first index of the smallest |component|
  int im = 0;
  float best = fabsf(d.x);
  if (fabsf(d.y) < best) { best = fabsf(d.y); im = 1; }
  if (fabsf(d.z) < best) im = 2;
  f3 n = f3{0.f, 0.f, 0.f};
  set_comp(n, im, 1.f);
  n = n - d * comp(d, im);
  return normalize(n);
}

// Quaternion product (nalgebra's Quaternion * Quaternion), coordinates (i, j, k, w).
__device__ __forceinline__ quat quat_mul(quat a, quat b) {
  quat r;
  r.w = ((a.w * b.w - a.i * b.i) - a.j * b.j) - a.k * b.k;
  r.i = ((a.w * b.i + a.i * b.w) + a.j * b.k) - a.k * b.j;
  r.j = ((a.w * b.j - a.i * b.k) + a.j * b.w) + a.k * b.i;
  r.k = ((a.w * b.k + a.i * b.j) - a.j * b.i) + a.k * b.w;
  return r;
}

// m1.inverse() * m2 (nalgebra Isometry::inverse then Isometry * Isometry): rotation q1^-1 q2,
// translation q1^-1 (-t1) + q1^-1 t2.
__device__ __forceinline__ iso iso_inv_mul(const iso& m1, const iso& m2) {
  const quat qi = quat{-m1.q.i, -m1.q.j, -m1.q.k, m1.q.w};
  iso r;
  const f3 t_inv = qrot(qi, -m1.t);
  r.t = t_inv + qrot(qi, m2.t);
  r.q = quat_mul(qi, m2.q);
  return r;
}

// |R| h of a cuboid rotated by q (isometry_ops.rs:43-45 through UnitQuaternion::to_rotation_matrix).
__device__ __forceinline__ f3 cuboid_world_half(quat q, float hx, float hy, float hz) {
  const float qi = q.i, qj = q.j, qk = q.k, qw = q.w;
  const float ww = qw * qw, ii = qi * qi, jj = qj * qj, kk = qk * qk;
  const float ij = qi * qj * 2.f, wk = qw * qk * 2.f, wj = qw * qj * 2.f;
  const float ik = qi * qk * 2.f, jk = qj * qk * 2.f, wi = qw * qi * 2.f;
  const float M[3][3] = {{ww + ii - jj - kk, ij - wk, wj + ik},
                         {wk + ij, ww - ii + jj - kk, jk - wi},
                         {ik - wj, wi + jk, ww - ii - jj + kk}};
  f3 ws;
  ws.x = (fabsf(M[0][0]) * hx + fabsf(M[0][1]) * hy) + fabsf(M[0][2]) * hz;
  ws.y = (fabsf(M[1][0]) * hx + fabsf(M[1][1]) * hy) + fabsf(M[1][2]) * hz;
  ws.z = (fabsf(M[2][0]) * hx + fabsf(M[2][1]) * hy) + fabsf(M[2][2]) * hz;
  return ws;
}

// Shape::aabb at m (aabb_ball.rs:8-13, aabb_cuboid.rs:7-14, aabb_support_map.rs:35-45,
// aabb_convex.rs:7-11 -> aabb_utils.rs:59-79), every lane of a group of G the same result (the
// hull's points split over the lanes; a min/max does not depend on the order).
template <int G>
__device__ __forceinline__ void shape_aabb_g(const ShapeD& g, const iso& m, const Pool& P, int gl, f3& mn, f3& mx) {
  if (g.type == NCX_BALL) {
    mn = m.t + f3{-g.a, -g.a, -g.a};
    mx = m.t + f3{g.a, g.a, g.a};
  } else if (g.type == NCX_CUBOID) {
    const f3 ws = cuboid_world_half(m.q, g.a, g.b, g.c);
    mn = m.t - ws;
    mx = m.t + ws;
  } else if (g.type == NCX_CAPSULE) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      f3 e = f3{0.f, 0.f, 0.f};
      set_comp(e, d, 1.f);
      set_comp(mx, d, comp(support<1>(g, m, e, P, 0), d));
      set_comp(e, d, -1.f);
      set_comp(mn, d, comp(support<1>(g, m, e, P, 0), d));
    }
  } else {
    f3 lo = f3{__builtin_inff(), __builtin_inff(), __builtin_inff()};
    f3 hi = -lo;
    for (uint32_t v = gl; v < g.cnt; v += G) {
      const f3 p = xform_pt(m, f3{P.x[g.off + v], P.y[g.off + v], P.z[g.off + v]});
      lo = f3{fminf(lo.x, p.x), fminf(lo.y, p.y), fminf(lo.z, p.z)};
      hi = f3{fmaxf(hi.x, p.x), fmaxf(hi.y, p.y), fmaxf(hi.z, p.z)};
    }
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) {
      lo.x = fminf(lo.x, __shfl_xor(lo.x, o, 64)); lo.y = fminf(lo.y, __shfl_xor(lo.y, o, 64));
      lo.z = fminf(lo.z, __shfl_xor(lo.z, o, 64));
      hi.x = fmaxf(hi.x, __shfl_xor(hi.x, o, 64)); hi.y = fmaxf(hi.y, __shfl_xor(hi.y, o, 64));
      hi.z = fmaxf(hi.z, __shfl_xor(hi.z, o, 64));
    }
    mn = lo;
    mx = hi;
  }
}
