// ncx_manifold.hip -- the narrow phase's contact-manifold generators, batched.
//
// One update of a pair's contact manifold (NarrowPhase::update_contact,
// src/pipeline/narrow_phase/narrow_phase.rs:56-80) with the generator DefaultContactDispatcher
// picks for Ball, Cuboid, Capsule and ConvexHull (default_contact_dispatcher.rs:27-97):
//   * Ball-Ball: one contact (ball_ball_manifold_generator.rs:27-66);
//   * Ball-polyhedron: the ball centre projected with its feature
//     (ball_convex_polyhedron_manifold_generator.rs:28-121);
//   * polyhedron-polyhedron: GJK/EPA from the cached last_gjk_dir, then the support faces (the
//     shapes overlap) or support features (within the prediction) clipped against each other
//     (convex_polyhedron_convex_polyhedron_manifold_generator.rs:83-167,
//     shape/convex_polygonal_feature3.rs:217-400);
//   * a capsule is its core segment with the prediction grown by its radius and a preprocessor
//     that moves its contact points out by the radius (capsule_shape_manifold_generator.rs,
//     capsule_capsule_manifold_generator.rs, shape/capsule.rs:86-127);
// and ContactManifold's distance-based tracking (query/contact/contact_manifold.rs:44-277):
// contacts within 0.02 of a tracked point (the first shape's local frame) replace it, the deeper
// one winning within one update.
//
// Two passes per call.  Pass 1 is the batched contact kernel in its manifold mode (capsules as
// segments, per-pair prediction, init_dir from each slot's last_gjk_dir): it gives every pair's
// GJK/EPA or closed-form contact and the new direction.  Pass 2 (k_manifold, one thread per pair)
// builds the polygonal features, clips them and updates the slot's tracked points.
//
// State per slot (one per interaction, owned by the caller's pair bookkeeping): last_gjk_dir
// (zero: None) and the tracking points of the last update's contacts, in cache order.
#include <cmath>
#include <thread>
#include <map>
#include "ncx_internal.h"
#include "ncx_gjk.h"
#include "ncx_contact.h"
#include "ncx_hullpoly.h"

#define MF_MAXV 16  // vertices of one polygonal feature (a hull face may have no more)
#define MF_MAXC NCX_MANIFOLD_MAX
static_assert(MF_MAXC <= 32, "a manifold's active entries are one 32-bit mask");

#define MF_FV 0u
#define MF_FE 1u
#define MF_FF 2u
#define MF_NONE 0xC0000000u  // FeatureId::Unknown
#define MF_ERR_FEATURE 1u    // a hull face with more than MF_MAXV vertices
#define MF_ERR_CACHE 2u      // more than MF_MAXC tracked contacts in one manifold
#define MF_ERR_SLOT 4u       // a slot index outside the state
#define MF_ERR_INDEX 8u      // a pair names a shape outside the set
#define MF_ERR_HULL 16u      // a pair names a hull with no polyhedral structure (try_new is None)

struct ncx_manifolds {
  ncx_ctx* ctx = nullptr;
  size_t nslots = 0;
  DevBuf dir;   // float4 per slot: last_gjk_dir (zero: None)
  DevBuf cnt;   // uint32 per slot: tracked points
  DevBuf pts;   // float4 per slot and entry: tracking points
  DevBuf idir, has, w1, w2, nrm, depth, fdir;  // the contact pass, per pair of the call
  DevBuf err;
  ContactTiers tiers;
};

__device__ __forceinline__ uint32_t mf_feat(uint32_t kind, uint32_t id) { return (kind << 30) | id; }

struct HullF {
  const uint4* sh;
  const float4* fn;
  const uint2* floop;
  const uint2* loop;
  const float4* edir;
  const uint2* ev;
  const uint2* vadj;
  const uint2* adj;
};

// ConvexPolygonalFeature (shape/convex_polygonal_feature3.rs:38-56) without its edge normals
// (only ConvexPolygonalFeature::project_point reads them).
struct PolyF {
  int nv, ne;
  f3 v[MF_MAXV];
  uint32_t vid[MF_MAXV];
  uint32_t eid[MF_MAXV];
  bool has_n;
  f3 n;
  uint32_t fid;
};

__device__ __forceinline__ void pf_clear(PolyF& f) {
  f.nv = 0;
  f.ne = 0;
  f.has_n = false;
  f.fid = MF_NONE;
}
__device__ __forceinline__ void pf_push(PolyF& f, f3 p, uint32_t id) {
  f.v[f.nv] = p;
  f.vid[f.nv] = id;
  f.nv++;
}
__device__ __forceinline__ void pf_transform(PolyF& f, const iso& m) {
  for (int i = 0; i < f.nv; ++i) f.v[i] = xform_pt(m, f.v[i]);
  if (f.has_n) f.n = xform_vec(m, f.n);
}
__device__ __forceinline__ int pf_nedges(const PolyF& f) { return f.nv == 1 ? 0 : (f.nv == 2 ? 1 : f.nv); }

// Cuboid::face (shape/cuboid.rs:185-276), 3D.
__device__ void cuboid_face(f3 h, uint32_t i, PolyF& out) {
  pf_clear(out);
  uint32_t i1;
  float sign;
  if (i < 3) { i1 = i; sign = 1.f; } else { i1 = i - 3; sign = -1.f; }
  const uint32_t i2 = (i1 + 1) % 3, i3 = (i1 + 2) % 3;
  const uint32_t ei2 = sign > 0.f ? i2 : i3, ei3 = sign > 0.f ? i3 : i2;
  const uint32_t mask_i2 = ~(1u << ei2), mask_i3 = ~(1u << ei3);
  f3 vertex = h;
  set_comp(vertex, i1, comp(vertex, i1) * sign);
  const uint32_t sbit = sign < 0.f ? 1u : 0u, msbit = sign < 0.f ? 0u : 1u;
  uint32_t vid = sbit << i1;
  pf_push(out, vertex, mf_feat(MF_FV, vid));
  out.eid[out.ne++] = mf_feat(MF_FE, ei2 | ((vid & mask_i2) << 2));
  set_comp(vertex, i2, -sign * comp(h, i2));
  set_comp(vertex, i3, sign * comp(h, i3));
  vid |= (msbit << i2) | (sbit << i3);
  pf_push(out, vertex, mf_feat(MF_FV, vid));
  out.eid[out.ne++] = mf_feat(MF_FE, ei3 | ((vid & mask_i3) << 2));
  set_comp(vertex, i2, -comp(h, i2));
  set_comp(vertex, i3, -comp(h, i3));
  vid |= (1u << i2) | (1u << i3);
  pf_push(out, vertex, mf_feat(MF_FV, vid));
  out.eid[out.ne++] = mf_feat(MF_FE, ei2 | ((vid & mask_i2) << 2));
  set_comp(vertex, i2, sign * comp(h, i2));
  set_comp(vertex, i3, -sign * comp(h, i3));
  vid = (sbit << i1) | (sbit << i2) | (msbit << i3);
  pf_push(out, vertex, mf_feat(MF_FV, vid));
  out.eid[out.ne++] = mf_feat(MF_FE, ei3 | ((vid & mask_i3) << 2));
  f3 nn = f3{0.f, 0.f, 0.f};
  set_comp(nn, i1, sign);
  out.has_n = true;
  out.n = nn;
  out.fid = mf_feat(MF_FF, sign > 0.f ? i1 : i1 + 3);
}

// ConvexHull::face (shape/convex.rs:443-460) of the hull of shape record g (faces at fbase).
__device__ void hull_face(const HullF& H, const Pool& P, const ShapeD& g, uint32_t fbase, uint32_t f, PolyF& out,
                          uint32_t& err) {
  pf_clear(out);
  const uint2 fl = H.floop[fbase + f];
  if (fl.y > MF_MAXV) {
    err |= MF_ERR_FEATURE;
    return;
  }
  for (uint32_t k = fl.x; k < fl.x + fl.y; ++k) {
    const uint2 le = H.loop[k];
    pf_push(out, f3{P.x[g.off + le.x], P.y[g.off + le.x], P.z[g.off + le.x]}, mf_feat(MF_FV, le.x));
    out.eid[out.ne++] = mf_feat(MF_FE, le.y);
  }
  const float4 n = H.fn[fbase + f];
  out.has_n = true;
  out.n = f3{n.x, n.y, n.z};
  out.fid = mf_feat(MF_FF, f);
}

// ConvexPolyhedron::support_face_toward: Cuboid (cuboid.rs:279-307), ConvexHull
// (convex.rs:487-509), Segment (segment.rs:291-304; the whole segment, whatever the direction).
__device__ void support_face_toward(const HullF& H, const Pool& P, const ShapeD& g, uint32_t si, const iso& m, f3 dir,
                                    PolyF& out, uint32_t& err) {
  pf_clear(out);
  if (g.type == NCX_CUBOID) {
    const f3 ld = inv_xform_vec(m, dir);
    int iamax = 0;
    float amax = fabsf(ld.x);
    if (fabsf(ld.y) > amax) { amax = fabsf(ld.y); iamax = 1; }
    if (fabsf(ld.z) > amax) iamax = 2;
    cuboid_face(f3{g.a, g.b, g.c}, comp(ld, iamax) > 0.f ? (uint32_t)iamax : (uint32_t)iamax + 3, out);
    pf_transform(out, m);
  } else if (g.type == NCX_CONVEX_HULL) {
    const f3 ld = inv_xform_vec(m, dir);
    const uint4 hs = H.sh[si];
    uint32_t best = 0;
    float4 n0 = H.fn[hs.x];
    float max_dot = dot(f3{n0.x, n0.y, n0.z}, ld);
    for (uint32_t f = 1; f < hs.y; ++f) {
      const float4 n = H.fn[hs.x + f];
      const float d = dot(f3{n.x, n.y, n.z}, ld);
      if (d > max_dot) { max_dot = d; best = f; }
    }
    hull_face(H, P, g, hs.x, best, out, err);
    pf_transform(out, m);
  } else {  // SHAPE_SEGMENT
    pf_push(out, f3{0.f, -g.a, 0.f}, mf_feat(MF_FV, 0));
    pf_push(out, f3{0.f, g.a, 0.f}, mf_feat(MF_FV, 1));
    out.eid[out.ne++] = mf_feat(MF_FE, 0);
    out.fid = mf_feat(MF_FE, 0);
    pf_transform(out, m);
  }
}

// ConvexHull::support_feature_id_toward_eps (convex.rs:388-415), local direction ld; seps/ceps:
// sine and cosine of the angle.
__device__ uint32_t hull_support_feature_id(const HullF& H, const Pool& P, const ShapeD& g, uint32_t si, f3 ld,
                                            float seps, float ceps) {
  // utils::point_cloud_support_point_id: the first maximum of p . dir
  uint32_t sv = 0;
  float best = dot(f3{P.x[g.off], P.y[g.off], P.z[g.off]}, ld);
  for (uint32_t i = 1; i < g.cnt; ++i) {
    const float d = dot(f3{P.x[g.off + i], P.y[g.off + i], P.z[g.off + i]}, ld);
    if (d > best) { best = d; sv = i; }
  }
  const uint4 hs = H.sh[si];
  const uint2 va = H.vadj[g.off + sv];
  for (uint32_t i = 0; i < va.y; ++i) {
    const uint32_t f = H.adj[va.x + i].x;
    const float4 n = H.fn[hs.x + f];
    if (dot(f3{n.x, n.y, n.z}, ld) >= ceps) return mf_feat(MF_FF, f);
  }
  for (uint32_t i = 0; i < va.y; ++i) {
    const uint32_t e = H.adj[va.x + i].y;
    const float4 d = H.edir[hs.z + e];
    if (fabsf(dot(f3{d.x, d.y, d.z}, ld)) <= seps) return mf_feat(MF_FE, e);
  }
  return mf_feat(MF_FV, sv);
}

// R::signum: +1 for +0 and positives, -1 for -0 and negatives, NaN for NaN.
__device__ __forceinline__ float rsignum(float x) { return x != x ? x : (__builtin_signbit(x) ? -1.f : 1.f); }

// ConvexPolyhedron::support_feature_toward: Cuboid (cuboid.rs:309-405), ConvexHull
// (convex.rs:511-540), Segment (segment.rs:306-344); cang/sang: cosine and sine of the angle.
__device__ void support_feature_toward(const HullF& H, const Pool& P, const ShapeD& g, uint32_t si, const iso& m,
                                       f3 dir, float cang, float sang, PolyF& out, uint32_t& err) {
  pf_clear(out);
  if (g.type == NCX_CUBOID) {
    const f3 h = f3{g.a, g.b, g.c};
    const f3 ld = inv_xform_vec(m, dir);
    f3 sp = h;
    uint32_t spid = 0;
    for (int i1 = 0; i1 < 3; ++i1) {
      const float sign = rsignum(comp(ld, i1));
      if (sign * comp(ld, i1) >= cang) {
        cuboid_face(h, sign > 0.f ? (uint32_t)i1 : (uint32_t)i1 + 3, out);
        pf_transform(out, m);
        return;
      } else if (sign < 0.f) {
        set_comp(sp, i1, comp(sp, i1) * sign);
        spid |= 1u << i1;
      }
    }
    for (int i = 0; i < 3; ++i) {
      const float sign = rsignum(comp(ld, i));
      if (sign * comp(ld, i) <= sang) {
        set_comp(sp, i, -comp(h, i));
        const f3 p1 = sp;
        set_comp(sp, i, comp(h, i));
        const f3 p2 = sp;
        const uint32_t p2id = spid & ~(1u << i);
        pf_push(out, xform_pt(m, p1), mf_feat(MF_FV, spid | (1u << i)));
        pf_push(out, xform_pt(m, p2), mf_feat(MF_FV, p2id));
        const uint32_t e = mf_feat(MF_FE, (uint32_t)i | (p2id << 2));
        out.eid[out.ne++] = e;
        out.fid = e;
        return;
      }
    }
    pf_push(out, xform_pt(m, sp), mf_feat(MF_FV, spid));
    out.fid = mf_feat(MF_FV, spid);
  } else if (g.type == NCX_CONVEX_HULL) {
    const f3 ld = inv_xform_vec(m, dir);
    const uint32_t f = hull_support_feature_id(H, P, g, si, ld, sang, cang);
    const uint32_t kind = f >> 30, id = f & 0x3FFFFFFFu;
    if (kind == MF_FV) {
      pf_push(out, f3{P.x[g.off + id], P.y[g.off + id], P.z[g.off + id]}, f);
      out.fid = f;
    } else if (kind == MF_FE) {
      const uint2 ev = H.ev[H.sh[si].z + id];
      pf_push(out, f3{P.x[g.off + ev.x], P.y[g.off + ev.x], P.z[g.off + ev.x]}, mf_feat(MF_FV, ev.x));
      pf_push(out, f3{P.x[g.off + ev.y], P.y[g.off + ev.y], P.z[g.off + ev.y]}, mf_feat(MF_FV, ev.y));
      out.fid = f;
      out.eid[out.ne++] = f;
    } else {
      hull_face(H, P, g, H.sh[si].x, id, out, err);
    }
    pf_transform(out, m);
  } else {  // SHAPE_SEGMENT, placed by m
    const f3 a = xform_pt(m, f3{0.f, -g.a, 0.f}), b = xform_pt(m, f3{0.f, g.a, 0.f});
    f3 sd;
    float nl;
    if (!try_normalize(b - a, NCX_FLT_EPS, &sd, &nl)) return;
    const float c = dot(dir, sd);
    if (c > sang) {
      out.fid = mf_feat(MF_FV, 1);
      pf_push(out, b, mf_feat(MF_FV, 1));
    } else if (c < -sang) {
      out.fid = mf_feat(MF_FV, 0);
      pf_push(out, a, mf_feat(MF_FV, 0));
    } else {
      pf_push(out, a, mf_feat(MF_FV, 0));
      pf_push(out, b, mf_feat(MF_FV, 1));
      out.eid[out.ne++] = mf_feat(MF_FE, 0);
      out.fid = mf_feat(MF_FE, 0);
    }
  }
}

// ------------------------------------------------------------------ the manifold's tracked contacts
struct MCache {
  uint32_t n;       // entries (old ones first, in cache order)
  uint32_t active;  // entries updated in this call
  float4* pts;      // the slot's tracking points
  size_t ob;        // first output entry of the pair
  float* w1;
  float* w2;
  float* nrm;
  float* depth;
  uint32_t* f1;
  uint32_t* f2;
};

struct CtOut {
  f3 w1, w2, n;
  float depth;
};

// capsule.rs:98-127 (CapsuleContactPreprocessor::process_contact)
__device__ __forceinline__ bool capsule_preprocess(float r, bool first, CtOut& c, uint32_t& f1, uint32_t& f2) {
  uint32_t& f = first ? f1 : f2;
  const uint32_t k = f >> 30;
  if (k == MF_FV) f = mf_feat(MF_FF, f & 0x3FFFFFFFu);
  else if (k == MF_FE) f = mf_feat(MF_FF, 2);
  else if (k == MF_FF) f = mf_feat(MF_FF, 2 + (f & 0x3FFFFFFFu));
  else return false;
  if (first) c.w1 = c.w1 + c.n * r;
  else c.w2 = c.w2 - c.n * r;
  c.depth = c.depth + r;
  return true;
}

// ContactManifold::push with DistanceBased(0.02) (contact_manifold.rs:168-223), after the
// contact preprocessors (cap > 0: the shape is a capsule's segment).
__device__ void manifold_push(MCache& M, CtOut c, uint32_t f1, uint32_t f2, f3 track, float cap1, float cap2,
                              uint32_t& err) {
  if (cap1 > 0.f && !capsule_preprocess(cap1, true, c, f1, f2)) return;
  if (cap2 > 0.f && !capsule_preprocess(cap2, false, c, f1, f2)) return;
  const float threshold = 0.02f;
  uint32_t closest = M.n;
  float cd = threshold * threshold;
  for (uint32_t i = 0; i < M.n; ++i) {
    const float4 p = M.pts[i];
    const float d = norm_sq(f3{p.x, p.y, p.z} - track);
    if (d < cd) { cd = d; closest = i; }
  }
  uint32_t j = closest;
  if (closest == M.n) {
    if (M.n >= MF_MAXC) {
      err |= MF_ERR_CACHE;
      return;
    }
    M.n++;
  } else if (M.active & (1u << j)) {
    if (c.depth <= M.depth[M.ob + j]) return;  // the contact already there is deeper
  }
  M.active |= 1u << j;
  M.pts[j] = make_float4(track.x, track.y, track.z, 0.f);
  const size_t o = M.ob + j;
  M.w1[3 * o] = c.w1.x; M.w1[3 * o + 1] = c.w1.y; M.w1[3 * o + 2] = c.w1.z;
  M.w2[3 * o] = c.w2.x; M.w2[3 * o + 1] = c.w2.y; M.w2[3 * o + 2] = c.w2.z;
  M.nrm[3 * o] = c.n.x; M.nrm[3 * o + 1] = c.n.y; M.nrm[3 * o + 2] = c.n.z;
  M.depth[o] = c.depth;
  M.f1[o] = f1;
  M.f2[o] = f2;
}

// ConvexPolygonalFeature::edge + Segment::direction: false when the edge is missing or degenerate.
__device__ __forceinline__ bool pf_edge_ok(const PolyF& F, uint32_t id) {
  for (int i1 = 0; i1 < F.nv && i1 < F.ne; ++i1)
    if (F.eid[i1] == id) {
      f3 d;
      float nl;
      return try_normalize(F.v[(i1 + 1) % F.nv] - F.v[i1], NCX_FLT_EPS, &d, &nl);
    }
  return false;
}

// ConvexPolygonalFeature::add_contact_to_manifold (convex_polygonal_feature3.rs:355-400).
__device__ void add_contact(const PolyF& F1, const PolyF& F2, const CtOut& c, uint32_t f1, uint32_t f2, const iso& m1,
                            float cap1, float cap2, MCache& M, uint32_t& err) {
  const f3 local1 = inv_xform_pt(m1, c.w1);
  if (f1 == MF_NONE || f2 == MF_NONE) return;
  if ((f1 >> 30) == MF_FE && !pf_edge_ok(F1, f1)) return;
  if ((f2 >> 30) == MF_FE && !pf_edge_ok(F2, f2)) return;
  manifold_push(M, c, f1, f2, local1, cap1, cap2, err);
}

// utils::point_in_poly2d (utils/point_in_poly2d.rs:5-26)
__device__ __forceinline__ bool point_in_poly2d(float px, float py, const float* xs, const float* ys, int n) {
  if (n == 0) return false;
  float sign = 0.f;
  for (int i1 = 0; i1 < n; ++i1) {
    const int i2 = (i1 + 1) % n;
    const float sx = xs[i2] - xs[i1], sy = ys[i2] - ys[i1];
    const float dx = px - xs[i1], dy = py - ys[i1];
    const float perp = dx * sy - dy * sx;
    if (sign == 0.f) sign = perp;
    else if (sign * perp < 0.f) return false;
  }
  return true;
}

// approx::ulps_eq!(a, b), default epsilon and max_ulps 4.
__device__ __forceinline__ bool ulps_eq_f(float a, float b) {
  if ((a > b ? a - b : b - a) <= NCX_FLT_EPS) return true;
  if (__builtin_signbit(a) != __builtin_signbit(b) || a != a || b != b) return false;
  const uint32_t ia = __float_as_uint(a), ib = __float_as_uint(b);
  return (ia <= ib ? ib - ia : ia - ib) <= 4u;
}
__device__ __forceinline__ float na_clamp(float v, float lo, float hi) { return v > lo ? (v < hi ? v : hi) : lo; }

// closest_points_segment_segment_with_locations_nD (closest_points_segment_segment.rs:44-141), 2D:
// true when both closest points are inside their segments, with their parameters.
__device__ bool seg_seg_2d_on_edges(float p1x, float p1y, float q1x, float q1y, float p2x, float p2y, float q2x,
                                    float q2y, float& s_out, float& t_out) {
  const float eps = NCX_FLT_EPS;
  const float d1x = q1x - p1x, d1y = q1y - p1y, d2x = q2x - p2x, d2y = q2y - p2y;
  const float rx = p1x - p2x, ry = p1y - p2y;
  const float a = d1x * d1x + d1y * d1y, e = d2x * d2x + d2y * d2y, f = d2x * rx + d2y * ry;
  float s, t;
  if (a <= eps && e <= eps) { s = 0.f; t = 0.f; }
  else if (a <= eps) { s = 0.f; t = na_clamp(f / e, 0.f, 1.f); }
  else {
    const float c = d1x * rx + d1y * ry;
    if (e <= eps) { t = 0.f; s = na_clamp(-c / a, 0.f, 1.f); }
    else {
      const float b = d1x * d2x + d1y * d2y;
      const float ae = a * e, bb = b * b;
      const float denom = ae - bb;
      const bool parallel = denom <= eps || ulps_eq_f(ae, bb);
      s = !parallel ? na_clamp((b * f - c * e) / denom, 0.f, 1.f) : 0.f;
      t = (b * s + f) / e;
      if (t < 0.f) { t = 0.f; s = na_clamp(-c / a, 0.f, 1.f); }
      else if (t > 1.f) { t = 1.f; s = na_clamp((b - c) / a, 0.f, 1.f); }
    }
  }
  s_out = s;
  t_out = t;
  return s != 0.f && s != 1.f && t != 0.f && t != 1.f;
}

// ConvexPolygonalFeature::clip (convex_polygonal_feature3.rs:217-352), each kept contact handed
// to add_contact in the reference's order; returns how many were kept.
__device__ int clip_and_add(const PolyF& S, const PolyF& O, f3 normal, float linear, const iso& m1, float cap1,
                            float cap2, MCache& M, uint32_t& err) {
  if (S.nv <= 2 && O.nv <= 2) return 0;
  // Vector::orthonormal_subspace_basis(&[normal]): a x normal, then a
  f3 a = fabsf(normal.x) > fabsf(normal.y) ? f3{normal.z, 0.f, -normal.x} : f3{0.f, -normal.z, normal.y};
  a = normalize(a);
  const f3 b0 = cross(a, normal), b1 = a;
  const f3 ref = S.v[0];
  float x1[MF_MAXV], y1[MF_MAXV], x2[MF_MAXV], y2[MF_MAXV];
  for (int i = 0; i < S.nv; ++i) { const f3 d = S.v[i] - ref; x1[i] = dot(b0, d); y1[i] = dot(b1, d); }
  for (int i = 0; i < O.nv; ++i) { const f3 d = O.v[i] - ref; x2[i] = dot(b0, d); y2[i] = dot(b1, d); }
  int kept = 0;
  auto emit = [&](f3 w1, f3 w2, uint32_t f1, uint32_t f2) {
    CtOut c;  // Contact::new_wo_depth (contact.rs:43-46)
    c.w1 = w1; c.w2 = w2; c.n = normal;
    c.depth = -dot(normal, w2 - w1);
    if (-c.depth <= linear) {
      kept++;
      add_contact(S, O, c, f1, f2, m1, cap1, cap2, M, err);
    }
  };
  if (O.nv > 2) {
    for (int i = 0; i < S.nv; ++i) {
      if (!point_in_poly2d(x1[i], y1[i], x2, y2, O.nv)) continue;
      const f3 origin = (ref + b0 * x1[i]) + b1 * y1[i];
      // query::line_toi_with_plane (query/ray/ray_plane.rs:9-23)
      const float denom = dot(O.n, normal);
      if (relative_eq_zero(denom)) continue;
      const float toi = dot(O.n, O.v[0] - origin) / denom;
      emit(S.v[i], origin + normal * toi, S.vid[i], O.fid);
    }
  }
  if (S.nv > 2) {
    for (int i = 0; i < O.nv; ++i) {
      if (!point_in_poly2d(x2[i], y2[i], x1, y1, S.nv)) continue;
      const f3 origin = (ref + b0 * x2[i]) + b1 * y2[i];
      const float denom = dot(S.n, normal);
      if (relative_eq_zero(denom)) continue;
      const float toi = dot(S.n, S.v[0] - origin) / denom;
      emit(origin + normal * toi, O.v[i], S.fid, O.vid[i]);
    }
  }
  const int ne1 = pf_nedges(S), ne2 = pf_nedges(O);
  for (int i1 = 0; i1 < ne1; ++i1) {
    const int j1 = (i1 + 1) % S.nv;
    for (int i2 = 0; i2 < ne2; ++i2) {
      const int j2 = (i2 + 1) % O.nv;
      float sp, tp;
      if (!seg_seg_2d_on_edges(x1[i1], y1[i1], x1[j1], y1[j1], x2[i2], y2[i2], x2[j2], y2[j2], sp, tp)) continue;
      // Segment::point_at(OnEdge([1 - s, s])) (shape/segment.rs:128-137)
      const f3 w1 = S.v[i1] * (1.f - sp) + S.v[j1] * sp;
      const f3 w2 = O.v[i2] * (1.f - tp) + O.v[j2] * tp;
      emit(w1, w2, S.eid[i1], O.eid[i2]);
    }
  }
  return kept;
}

struct MArgs {
  const ncx_shape* shapes;
  uint32_t nshapes;
  Pool P;
  HullF H;
  const ncx_pose* poses;
  const uint32_t* ia;
  const uint32_t* ib;
  const uint32_t* slots;
  size_t npairs, nslots;
  float linear, c1, s1, c2, s2, s1deg, c1deg;
  const uint8_t* has;
  const float* w1;
  const float* w2;
  const float* nrm;
  const float* depth;
  const float* fdir;
  float4* st_dir;
  uint32_t* st_cnt;
  float4* st_pts;
  uint32_t* count;
  float* ow1;
  float* ow2;
  float* onrm;
  float* odepth;
  uint32_t* of1;
  uint32_t* of2;
  uint32_t* err;
};

__global__ void k_mf_gather(const uint32_t* __restrict__ slots, size_t n, size_t nslots,
                            const float4* __restrict__ st_dir, float* __restrict__ idir) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = slots[k];
  const float4 d = s < nslots ? st_dir[s] : make_float4(0.f, 0.f, 0.f, 0.f);
  idir[3 * k] = d.x;
  idir[3 * k + 1] = d.y;
  idir[3 * k + 2] = d.z;
}

__global__ __launch_bounds__(64) void k_manifold(MArgs a) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.npairs) return;
  uint32_t err = 0;
  const uint32_t slot = a.slots[k];
  const uint32_t i1 = a.ia[k], i2 = a.ib[k];
  if (slot >= a.nslots || i1 >= a.nshapes || i2 >= a.nshapes) {
    atomicOr(a.err, slot >= a.nslots ? MF_ERR_SLOT : MF_ERR_INDEX);
    a.count[k] = 0;
    return;
  }
  const ShapeD g1 = load_shape(a.shapes, i1), g2 = load_shape(a.shapes, i2);
  if ((g1.type == NCX_CONVEX_HULL && a.H.sh[i1].w) || (g2.type == NCX_CONVEX_HULL && a.H.sh[i2].w)) {
    atomicOr(a.err, MF_ERR_HULL);
    a.count[k] = 0;
    return;
  }
  const iso m1 = load_pose(a.poses, i1), m2 = load_pose(a.poses, i2);
  MCache M;
  M.n = a.st_cnt[slot];  // save_cache_and_clear: the last update's contacts are the old entries
  M.active = 0;
  M.pts = a.st_pts + (size_t)slot * MF_MAXC;
  M.ob = k * MF_MAXC;
  M.w1 = a.ow1; M.w2 = a.ow2; M.nrm = a.onrm; M.depth = a.odepth; M.f1 = a.of1; M.f2 = a.of2;
  ShapeD s1 = g1, s2 = g2;
  float cap1 = 0.f, cap2 = 0.f, lin = a.linear;
  if (g1.type == NCX_CAPSULE) { s1.type = SHAPE_SEGMENT; cap1 = g1.b; lin = lin + g1.b; }
  if (g2.type == NCX_CAPSULE) { s2.type = SHAPE_SEGMENT; cap2 = g2.b; lin = lin + g2.b; }
  const bool has = a.has[k] == 1;
  CtOut c;
  c.w1 = f3{a.w1[3 * k], a.w1[3 * k + 1], a.w1[3 * k + 2]};
  c.w2 = f3{a.w2[3 * k], a.w2[3 * k + 1], a.w2[3 * k + 2]};
  c.n = f3{a.nrm[3 * k], a.nrm[3 * k + 1], a.nrm[3 * k + 2]};
  c.depth = a.depth[k];
  const uint32_t ball = mf_feat(MF_FF, 0);
  if (s1.type == NCX_BALL && s2.type == NCX_BALL) {
    if (has) manifold_push(M, c, ball, ball, f3{0.f, 0.f, 0.f}, 0.f, 0.f, err);
  } else if (s1.type == NCX_BALL || s2.type == NCX_BALL) {
    if (has) {
      // the polyhedron's feature at the ball centre's projection (project_point_with_feature)
      const bool flip = s2.type == NCX_BALL;
      const ShapeD& gp = flip ? s1 : s2;
      const uint32_t sp = flip ? i1 : i2;
      const iso& mp = flip ? m1 : m2;
      const f3 center = flip ? m2.t : m1.t;
      const f3 proj = flip ? c.w1 : c.w2;
      const f3 nball = flip ? -c.n : c.n;  // the normal as the ball-first generator set it
      uint32_t f = MF_NONE;
      if (gp.type == NCX_CUBOID) {
        bool inside;
        int fk, fi;
        (void)project_cuboid(mp, f3{gp.a, gp.b, gp.c}, center, inside, fk, fi);
        f = fk == 0 ? mf_feat(MF_FF, fi) : fk == 1 ? mf_feat(MF_FE, fi) : fk == 2 ? mf_feat(MF_FV, fi) : MF_NONE;
      } else if (gp.type == SHAPE_SEGMENT) {
        int kind, id;
        bool inside;
        (void)proj_segment_at(mp, f3{0.f, -gp.a, 0.f}, f3{0.f, gp.a, 0.f}, center, kind, id, inside);
        f = kind == LOC_VERTEX ? mf_feat(MF_FV, id) : mf_feat(MF_FE, 0);
      } else {
        // ConvexHull::project_point_with_feature (point_support_map.rs:96-118); inside <=> the
        // normal points from the projection back to the centre
        const bool inside = dot(nball, proj - center) < 0.f;
        const f3 dpt = center - proj;
        const f3 ld = inside ? inv_xform_vec(mp, -dpt) : inv_xform_vec(mp, dpt);
        f3 u;
        float nl;
        if (try_normalize(ld, NCX_FLT_EPS, &u, &nl)) f = hull_support_feature_id(a.H, a.P, gp, sp, u, a.s1deg, a.c1deg);
      }
      if (f != MF_NONE) {
        if (!flip) manifold_push(M, c, ball, f, f3{0.f, 0.f, 0.f}, 0.f, cap2, err);
        else manifold_push(M, c, f, ball, f3{0.f, 0.f, 0.f}, cap1, 0.f, err);
      }
    }
  } else {
    // ConvexPolyhedronConvexPolyhedronManifoldGenerator: the contact pass ran GJK/EPA from the
    // slot's last_gjk_dir; keep its result direction (unchanged when it carried none)
    a.st_dir[slot] = make_float4(a.fdir[3 * k], a.fdir[3 * k + 1], a.fdir[3 * k + 2], 0.f);
    if (has) {
      PolyF F1, F2;
      if (c.depth > 0.f) {
        support_face_toward(a.H, a.P, s1, i1, m1, c.n, F1, err);
        support_face_toward(a.H, a.P, s2, i2, m2, -c.n, F2, err);
      } else {
        support_feature_toward(a.H, a.P, s1, i1, m1, c.n, a.c1, a.s1, F1, err);
        support_feature_toward(a.H, a.P, s2, i2, m2, -c.n, a.c2, a.s2, F2, err);
      }
      if (clip_and_add(F1, F2, c.n, lin, m1, cap1, cap2, M, err) == 0)
        add_contact(F1, F2, c, F1.fid, F2.fid, m1, cap1, cap2, M, err);
    }
  }
  // the contacts of this update, in cache order; their tracking points stay for the next one
  uint32_t out = 0;
  for (uint32_t j = 0; j < M.n; ++j) {
    if (!(M.active & (1u << j))) continue;
    if (out != j) {
      M.pts[out] = M.pts[j];
      const size_t s = M.ob + j, d = M.ob + out;
      for (int q = 0; q < 3; ++q) {
        a.ow1[3 * d + q] = a.ow1[3 * s + q];
        a.ow2[3 * d + q] = a.ow2[3 * s + q];
        a.onrm[3 * d + q] = a.onrm[3 * s + q];
      }
      a.odepth[d] = a.odepth[s];
      a.of1[d] = a.of1[s];
      a.of2[d] = a.of2[s];
    }
    out++;
  }
  a.st_cnt[slot] = out;
  a.count[k] = out;
  if (err) atomicOr(a.err, err);
}

// ------------------------------------------------------------------ host side
// ConvexHull::try_new on every distinct hull of the set (host threads), uploaded once.
static int build_hull_features(ncx_shapes* S) {
  if (S->hf_ready) return NCX_OK;
  ncx_ctx* ctx = S->ctx;
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  std::vector<ncx_shape> rec(S->n);
  std::vector<float> X(S->npts), Y(S->npts), Z(S->npts);
  if (S->n) NCX_HIP_CHECK(hipMemcpy(rec.data(), S->shapes.p, S->n * sizeof(ncx_shape), hipMemcpyDeviceToHost));
  if (S->npts) {
    NCX_HIP_CHECK(hipMemcpy(X.data(), S->hx.p, S->npts * 4, hipMemcpyDeviceToHost));
    NCX_HIP_CHECK(hipMemcpy(Y.data(), S->hy.p, S->npts * 4, hipMemcpyDeviceToHost));
    NCX_HIP_CHECK(hipMemcpy(Z.data(), S->hz.p, S->npts * 4, hipMemcpyDeviceToHost));
  }
  // one structure per distinct (offset, count) range
  std::map<std::pair<uint32_t, uint32_t>, size_t> key;
  std::vector<std::pair<uint32_t, uint32_t>> ranges;
  std::vector<size_t> which(S->n, (size_t)-1);
  for (size_t i = 0; i < S->n; ++i) {
    if (rec[i].type != NCX_CONVEX_HULL) continue;
    auto kk = std::make_pair(rec[i].hull_offset, rec[i].hull_count);
    auto it = key.find(kk);
    if (it == key.end()) {
      it = key.emplace(kk, ranges.size()).first;
      ranges.push_back(kk);
    }
    which[i] = it->second;
  }
  // per-vertex adjacency is indexed by pool point: distinct ranges must not share points
  {
    std::vector<uint32_t> owner(S->npts, 0xFFFFFFFFu);
    for (size_t r = 0; r < ranges.size(); ++r)
      for (uint32_t i = 0; i < ranges[r].second; ++i) {
        uint32_t& o = owner[ranges[r].first + i];
        NCX_REQUIRE(o == 0xFFFFFFFFu, "contact manifolds: two hulls overlap in the point pool");
        o = (uint32_t)r;
      }
  }
  std::vector<HullPoly> polys(ranges.size());
  std::vector<uint8_t> ok(ranges.size(), 0);
  unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nth; ++t)
    th.emplace_back([&, t]() {
      std::vector<float> xyz;
      std::vector<uint32_t> tris;
      for (size_t r = t; r < ranges.size(); r += nth) {
        const uint32_t off = ranges[r].first, cnt = ranges[r].second;
        xyz.resize(3 * (size_t)cnt);
        for (uint32_t i = 0; i < cnt; ++i) {
          xyz[3 * i] = X[off + i];
          xyz[3 * i + 1] = Y[off + i];
          xyz[3 * i + 2] = Z[off + i];
        }
        try {
          if (!hull_triangulate(xyz.data(), cnt, tris)) continue;
          ok[r] = hull_poly_build(X.data() + off, Y.data() + off, Z.data() + off, cnt, tris.data(),
                                  (uint32_t)(tris.size() / 3), polys[r]) ? 1 : 0;
        } catch (...) {
          ok[r] = 0;
        }
      }
    });
  for (auto& x : th) x.join();
  // a hull whose try_new fails is marked; a pair naming it fails the call (the reference has no
  // such ConvexHull: ConvexHull::try_from_points returns None)
  for (size_t r = 0; r < ranges.size(); ++r)
    if (!ok[r]) polys[r] = HullPoly();
  // concatenate: faces, loops, edges, adjacency
  std::vector<size_t> fb(ranges.size()), lb(ranges.size()), eb(ranges.size()), ab(ranges.size());
  size_t nf = 0, nl = 0, ne = 0, na = 0;
  for (size_t r = 0; r < ranges.size(); ++r) {
    fb[r] = nf; lb[r] = nl; eb[r] = ne; ab[r] = na;
    nf += polys[r].ffirst.size();
    nl += polys[r].lvert.size();
    ne += polys[r].ev0.size();
    na += polys[r].vface.size();
  }
  NCX_REQUIRE(nf < 0xFFFFFFFFull && nl < 0xFFFFFFFFull && ne < 0xFFFFFFFFull && na < 0xFFFFFFFFull,
              "contact manifolds: hull structure too large");
  std::vector<uint32_t> hsh(4 * std::max<size_t>(S->n, 1), 0);
  std::vector<float> hfn(4 * std::max<size_t>(nf, 1), 0.f), hed(4 * std::max<size_t>(ne, 1), 0.f);
  std::vector<uint32_t> hfl(2 * std::max<size_t>(nf, 1), 0), hlp(2 * std::max<size_t>(nl, 1), 0),
      hev(2 * std::max<size_t>(ne, 1), 0), hva(2 * std::max<size_t>(S->npts, 1), 0), had(2 * std::max<size_t>(na, 1), 0);
  for (size_t r = 0; r < ranges.size(); ++r) {
    const HullPoly& H = polys[r];
    for (size_t f = 0; f < H.ffirst.size(); ++f) {
      hfn[4 * (fb[r] + f)] = H.fnx[f];
      hfn[4 * (fb[r] + f) + 1] = H.fny[f];
      hfn[4 * (fb[r] + f) + 2] = H.fnz[f];
      hfl[2 * (fb[r] + f)] = (uint32_t)(lb[r] + H.ffirst[f]);
      hfl[2 * (fb[r] + f) + 1] = H.fnum[f];
    }
    for (size_t k = 0; k < H.lvert.size(); ++k) {
      hlp[2 * (lb[r] + k)] = H.lvert[k];
      hlp[2 * (lb[r] + k) + 1] = H.ledge[k];
    }
    for (size_t e = 0; e < H.ev0.size(); ++e) {
      hed[4 * (eb[r] + e)] = H.edx[e];
      hed[4 * (eb[r] + e) + 1] = H.edy[e];
      hed[4 * (eb[r] + e) + 2] = H.edz[e];
      hev[2 * (eb[r] + e)] = H.ev0[e];
      hev[2 * (eb[r] + e) + 1] = H.ev1[e];
    }
    const uint32_t off = ranges[r].first;
    for (size_t v = 0; v < H.vfirst.size(); ++v) {
      hva[2 * (off + v)] = (uint32_t)(ab[r] + H.vfirst[v]);
      hva[2 * (off + v) + 1] = H.vnum[v];
    }
    for (size_t k = 0; k < H.vface.size(); ++k) {
      had[2 * (ab[r] + k)] = H.vface[k];
      had[2 * (ab[r] + k) + 1] = H.vedge[k];
    }
  }
  for (size_t i = 0; i < S->n; ++i) {
    if (which[i] == (size_t)-1) continue;
    const size_t r = which[i];
    hsh[4 * i] = (uint32_t)fb[r];
    hsh[4 * i + 1] = (uint32_t)polys[r].ffirst.size();
    hsh[4 * i + 2] = (uint32_t)eb[r];
    hsh[4 * i + 3] = ok[r] ? 0u : 1u;
  }
  int rc;
  auto up = [&](DevBuf& b, const void* src, size_t bytes) -> int {
    int e = b.reserve(bytes);
    if (e) return e;
    NCX_HIP_CHECK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    return NCX_OK;
  };
  if ((rc = up(S->hf_shape, hsh.data(), hsh.size() * 4)) || (rc = up(S->hf_fn, hfn.data(), hfn.size() * 4)) ||
      (rc = up(S->hf_floop, hfl.data(), hfl.size() * 4)) || (rc = up(S->hf_loop, hlp.data(), hlp.size() * 4)) ||
      (rc = up(S->hf_edir, hed.data(), hed.size() * 4)) || (rc = up(S->hf_ev, hev.data(), hev.size() * 4)) ||
      (rc = up(S->hf_vadj, hva.data(), hva.size() * 4)) || (rc = up(S->hf_adj, had.data(), had.size() * 4)))
    return rc;
  S->hf_ready = true;
  return NCX_OK;
}

extern "C" int ncx_manifolds_create(ncx_ctx* ctx, size_t nslots, ncx_manifolds** out) {
  NCX_REQUIRE(ctx && out && nslots > 0, "ncx_manifolds_create: bad argument");
  NCX_REQUIRE(nslots < ((size_t)1 << 32) / MF_MAXC, "ncx_manifolds_create: too many slots");
  ncx_manifolds* m = new ncx_manifolds;
  m->ctx = ctx;
  m->nslots = nslots;
  int rc;
  if ((rc = m->dir.reserve(nslots * 16)) || (rc = m->cnt.reserve(nslots * 4)) ||
      (rc = m->pts.reserve(nslots * MF_MAXC * 16)) || (rc = m->err.reserve(16))) {
    ncx_manifolds_destroy(m);
    return rc;
  }
  NCX_HIP_CHECK(hipMemsetAsync(m->dir.p, 0, nslots * 16, ctx->stream));
  NCX_HIP_CHECK(hipMemsetAsync(m->cnt.p, 0, nslots * 4, ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  *out = m;
  return NCX_OK;
}

extern "C" int ncx_manifolds_destroy(ncx_manifolds* m) {
  if (!m) return NCX_OK;
  m->dir.release(); m->cnt.release(); m->pts.release();
  m->idir.release(); m->has.release(); m->w1.release(); m->w2.release(); m->nrm.release();
  m->depth.release(); m->fdir.release(); m->err.release();
  contact_tiers_release(m->tiers);
  delete m;
  return NCX_OK;
}

__global__ void k_mf_reset(const uint32_t* __restrict__ slots, size_t n, size_t nslots, float4* __restrict__ dir,
                           uint32_t* __restrict__ cnt, uint32_t* __restrict__ err) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = slots[k];
  if (s >= nslots) {
    atomicOr(err, MF_ERR_SLOT);
    return;
  }
  dir[s] = make_float4(0.f, 0.f, 0.f, 0.f);
  cnt[s] = 0;
}

static int mf_check_err(ncx_manifolds* m) {
  uint32_t e = 0;
  NCX_HIP_CHECK(hipMemcpyAsync(&e, m->err.p, 4, hipMemcpyDeviceToHost, m->ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
  if (e & MF_ERR_SLOT) { ncx_set_error("contact manifolds: a slot index is outside the state"); return NCX_E_INVALID; }
  if (e & MF_ERR_INDEX) { ncx_set_error("contact manifolds: a pair index is not a shape of the set"); return NCX_E_INVALID; }
  if (e & MF_ERR_HULL) {
    ncx_set_error("contact manifolds: a pair names a hull with no polyhedral structure (ConvexHull::try_new "
                  "fails on its triangles: merged faces break Euler's characteristic)");
    return NCX_E_INVALID;
  }
  if (e & MF_ERR_FEATURE) {
    ncx_set_error("contact manifolds: a hull face has more than 16 vertices");
    return NCX_E_CAPACITY;
  }
  if (e & MF_ERR_CACHE) {
    ncx_set_error("contact manifolds: a manifold tracks more than NCX_MANIFOLD_MAX contacts");
    return NCX_E_CAPACITY;
  }
  return NCX_OK;
}

extern "C" int ncx_manifolds_reset(ncx_manifolds* m, const uint32_t* slots, size_t n) {
  NCX_REQUIRE(m && (slots || n == 0), "ncx_manifolds_reset: null argument");
  if (n == 0) return NCX_OK;
  int rc;
  if ((rc = fill_u32(m->ctx, m->err.as<uint32_t>(), 1, 0u))) return rc;
  hipLaunchKernelGGL(k_mf_reset, dim3(ceil_div(n, 256)), dim3(256), 0, m->ctx->stream, slots, n, m->nslots,
                     m->dir.as<float4>(), m->cnt.as<uint32_t>(), m->err.as<uint32_t>());
  NCX_HIP_CHECK(hipGetLastError());
  return mf_check_err(m);
}

extern "C" int ncx_manifolds_dirs(ncx_manifolds* m, const uint32_t* slots, size_t n, float* dirs) {
  NCX_REQUIRE(m && ((slots && dirs) || n == 0), "ncx_manifolds_dirs: null argument");
  if (n == 0) return NCX_OK;
  hipLaunchKernelGGL(k_mf_gather, dim3(ceil_div(n, 256)), dim3(256), 0, m->ctx->stream, slots, n, m->nslots,
                     m->dir.as<float4>(), dirs);
  NCX_HIP_CHECK(hipGetLastError());
  NCX_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
  return NCX_OK;
}

extern "C" int ncx_contact_manifolds(ncx_manifolds* m, ncx_shapes* s, const ncx_pose* poses, const uint32_t* ia,
                                     const uint32_t* ib, const uint32_t* slots, size_t npairs, float linear,
                                     float angular1, float angular2, uint32_t* count, float* world1, float* world2,
                                     float* normal, float* depth, uint32_t* feature1, uint32_t* feature2) {
  NCX_REQUIRE(m && s && poses && ia && ib && slots && count && world1 && world2 && normal && depth && feature1 &&
                  feature2,
              "ncx_contact_manifolds: null argument");
  NCX_REQUIRE(s->ctx == m->ctx, "ncx_contact_manifolds: the shape set and the state belong to different contexts");
  NCX_REQUIRE(!s->tri, "ncx_contact_manifolds: a composite shape's part set is not a shape set of its own");
  NCX_REQUIRE(npairs < ((size_t)1 << 31) / MF_MAXC, "ncx_contact_manifolds: too many pairs");
  if (npairs == 0) return NCX_OK;
  ncx_ctx* ctx = m->ctx;
  int rc;
  if ((rc = build_hull_features(s))) return rc;
  if ((rc = m->idir.reserve(npairs * 12)) || (rc = m->has.reserve(npairs)) || (rc = m->w1.reserve(npairs * 12)) ||
      (rc = m->w2.reserve(npairs * 12)) || (rc = m->nrm.reserve(npairs * 12)) || (rc = m->depth.reserve(npairs * 4)) ||
      (rc = m->fdir.reserve(npairs * 12)) || (rc = contact_tiers_reserve(m->tiers, npairs)))
    return rc;
  if ((rc = fill_u32(ctx, m->err.as<uint32_t>(), 1, 0u))) return rc;
  const unsigned nb = ceil_div(npairs, 256);
  hipLaunchKernelGGL(k_mf_gather, dim3(nb), dim3(256), 0, ctx->stream, slots, npairs, m->nslots, m->dir.as<float4>(),
                     m->idir.as<float>());
  NCX_HIP_CHECK(hipGetLastError());
  // pass 1: GJK/EPA or the closed forms, capsules as segments, from each slot's last_gjk_dir
  if ((rc = contacts_launch_device_count(s, m->tiers, ctx, poses, ia, ib, nullptr, npairs, linear, m->idir.as<float>(),
                                         10000, m->has.as<uint8_t>(), m->w1.as<float>(), m->w2.as<float>(),
                                         m->nrm.as<float>(), m->depth.as<float>(), nullptr, true, m->fdir.as<float>(),
                                         true)))
    return rc;
  if ((rc = contacts_resolve_overflow(m->tiers, ctx, nullptr))) return rc;
  // pass 2: features, clipping and tracking
  MArgs a = {};
  a.shapes = s->shapes.as<ncx_shape>();
  a.nshapes = (uint32_t)s->n;
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>(), nullptr};
  a.H = HullF{s->hf_shape.as<uint4>(), s->hf_fn.as<float4>(), s->hf_floop.as<uint2>(), s->hf_loop.as<uint2>(),
              s->hf_edir.as<float4>(), s->hf_ev.as<uint2>(), s->hf_vadj.as<uint2>(), s->hf_adj.as<uint2>()};
  a.poses = poses;
  a.ia = ia;
  a.ib = ib;
  a.slots = slots;
  a.npairs = npairs;
  a.nslots = m->nslots;
  a.linear = linear;
  // cos / sin of the prediction's angles (f32, as R::cos / R::sin), and of one degree
  // (ConvexHull::support_feature_id_toward, convex.rs:542-545)
  a.c1 = std::cos(angular1);
  a.s1 = std::sin(angular1);
  a.c2 = std::cos(angular2);
  a.s2 = std::sin(angular2);
  const float one_degree = (float)(3.14159265358979323846 / 180.0);
  a.s1deg = std::sin(one_degree);
  a.c1deg = std::cos(one_degree);
  a.has = m->has.as<uint8_t>();
  a.w1 = m->w1.as<float>();
  a.w2 = m->w2.as<float>();
  a.nrm = m->nrm.as<float>();
  a.depth = m->depth.as<float>();
  a.fdir = m->fdir.as<float>();
  a.st_dir = m->dir.as<float4>();
  a.st_cnt = m->cnt.as<uint32_t>();
  a.st_pts = m->pts.as<float4>();
  a.count = count;
  a.ow1 = world1;
  a.ow2 = world2;
  a.onrm = normal;
  a.odepth = depth;
  a.of1 = feature1;
  a.of2 = feature2;
  a.err = m->err.as<uint32_t>();
  hipLaunchKernelGGL(k_manifold, dim3(ceil_div(npairs, 64)), dim3(64), 0, ctx->stream, a);
  NCX_HIP_CHECK(hipGetLastError());
  return mf_check_err(m);
}
