// ncx_contact.h -- device EPA (LDS or device-memory arenas) and the per-pair contact dispatch,
// shared by the batched contact kernels and the composite-shape kernels.
//   EPA      src/query/algorithms/epa3.rs:171-454
//   dispatch src/query/contact/contact_shape_shape.rs:10-48, contact_ball_ball.rs:8-38,
//            contact_ball_convex_polyhedron.rs:11-77, contact_support_map_support_map.rs:38-88
#pragma once
#include "ncx_gjk.h"

// ------------------------------------------------------------------ EPA arena
// A polytope vertex keeps the two support points; its CSO point is recomputed as o1 - o2,
// which is how it was formed (CSOPoint::new), so the bits are the same.  Face and vertex ids are
// 16-bit (arena capacities stay below 65536).
struct VertA {
  float o1x, o1y, o1z, o2x, o2y, o2z;
};
struct FaceD {
  uint16_t pts[3];
  uint16_t adj[3];
  uint16_t deleted, pad;
};
struct HeapE {
  float neg_dist;
  uint32_t id;
};
struct SilE {
  uint16_t face, opp;
};

struct Arena {
  VertA* verts;
  FaceD* faces;
  HeapE* heap;
  SilE* sil;
  SilE* stack;
  uint32_t vcap, fcap, scap, stcap;  // vertices, faces (and heap entries), silhouette, search stack
};

__host__ __device__ constexpr size_t arena_bytes(uint32_t vcap, uint32_t fcap, uint32_t scap, uint32_t stcap) {
  return ((size_t)vcap * sizeof(VertA) + (size_t)fcap * (sizeof(FaceD) + sizeof(HeapE)) +
          (size_t)(scap + stcap) * sizeof(SilE) + 15) & ~(size_t)15;
}
// Device-memory tiers: silhouette and search stack sized by the face capacity.
__host__ __device__ constexpr size_t arena_bytes(uint32_t vcap, uint32_t fcap) {
  return arena_bytes(vcap, fcap, fcap, 2 * fcap);
}

__device__ __forceinline__ Arena arena_at(uint8_t* base, size_t slot, uint32_t vcap, uint32_t fcap, uint32_t scap,
                                          uint32_t stcap) {
  uint8_t* p = base + slot * arena_bytes(vcap, fcap, scap, stcap);
  Arena a;
  a.verts = (VertA*)p;
  p += (size_t)vcap * sizeof(VertA);
  a.faces = (FaceD*)p;
  p += (size_t)fcap * sizeof(FaceD);
  a.heap = (HeapE*)p;
  p += (size_t)fcap * sizeof(HeapE);
  a.sil = (SilE*)p;
  p += (size_t)scap * sizeof(SilE);
  a.stack = (SilE*)p;
  a.vcap = vcap;
  a.fcap = fcap;
  a.scap = scap;
  a.stcap = stcap;
  return a;
}

// Field-wise vertex copies between registers and the arena (no aggregate copies through
// private memory).
__device__ __forceinline__ CSO vget(const Arena& A, uint32_t i) {
  const float* f = (const float*)(A.verts + i);
  CSO c;
  c.o1 = f3{f[0], f[1], f[2]};
  c.o2 = f3{f[3], f[4], f[5]};
  c.p = c.o1 - c.o2;
  return c;
}
__device__ __forceinline__ f3 vpt(const Arena& A, uint32_t i) {
  const float* f = (const float*)(A.verts + i);
  return f3{f[0], f[1], f[2]} - f3{f[3], f[4], f[5]};
}
__device__ __forceinline__ void vset(const Arena& A, uint32_t i, const CSO& c) {
  float* f = (float*)(A.verts + i);
  f[0] = c.o1.x; f[1] = c.o1.y; f[2] = c.o1.z;
  f[3] = c.o2.x; f[4] = c.o2.y; f[5] = c.o2.z;
}

// Rust std BinaryHeap (max-heap on neg_dist): push = sift_up; pop = swap with the
// last element, sift_down_to_bottom, sift_up.  Element order is PartialOrd::le.
__device__ __forceinline__ void heap_sift_up(HeapE* h, uint32_t pos) {
  HeapE e = h[pos];
  while (pos > 0) {
    uint32_t parent = (pos - 1) / 2;
    if (e.neg_dist <= h[parent].neg_dist) break;
    h[pos] = h[parent];
    pos = parent;
  }
  h[pos] = e;
}
// heap_sift_up with the group's lanes: lane j reads the element's j-th ancestor (0 = the parent;
// in 1-based positions ancestor j of k is k >> (j + 1)), the sequential loop's break test
// `e <= ancestor` is evaluated on every ancestor at once, and the first ancestor where it holds
// (nearest first) is where the climb stops.  The ancestors below it each move down one level
// along the path and the element takes the last one's slot, which is what the loop leaves.
// Needs as many lanes as the heap has levels above pos.
template <int G, bool GMEM>
__device__ __forceinline__ void heap_sift_up_par(HeapE* h, uint32_t pos, int gl) {
  const HeapE e = h[pos];
  const uint32_t k = pos + 1;
  const uint32_t depth = 31u - (uint32_t)__builtin_clz(k);  // ancestors of pos
  const bool have = (uint32_t)gl < depth;
  const uint32_t anc = have ? (k >> (gl + 1)) - 1u : 0u;
  HeapE a = HeapE{0.f, 0u};
  if (have) a = h[anc];
  const bool stop = have && e.neg_dist <= a.neg_dist;
  const int gbase = __lane_id() & ~(G - 1);
  const uint64_t sm = (__ballot(stop) >> gbase) & (G == 64 ? ~0ull : ((1ull << (G & 63)) - 1ull));
  const uint32_t s = sm ? (uint32_t)__builtin_ctzll(sm) : depth;
  __builtin_amdgcn_wave_barrier();
  if ((uint32_t)gl < s) h[gl == 0 ? pos : (k >> gl) - 1u] = a;
  __builtin_amdgcn_wave_barrier();
  if (gl == 0) h[s == 0 ? pos : (k >> s) - 1u] = e;
  __builtin_amdgcn_wave_barrier();
  // A device-memory heap: the other lanes' stores must have landed before anyone reads the heap.
  if constexpr (GMEM) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

template <int G, int PAR>
__device__ __forceinline__ bool heap_pop(HeapE* h, uint32_t& len, HeapE& out, int gl) {
  if (len == 0) return false;
  len -= 1;
  HeapE item = h[len];
  if (len > 0) {
    HeapE top = h[0];
    h[0] = item;
    item = top;
    uint32_t end = len, pos = 0, child = 1;
    HeapE e = h[0];
    while (end >= 2 && child <= end - 2) {
      child += (h[child].neg_dist <= h[child + 1].neg_dist) ? 1u : 0u;
      h[pos] = h[child];
      pos = child;
      child = 2 * pos + 1;
    }
    if (child == end - 1) { h[pos] = h[child]; pos = child; }
    h[pos] = e;
    if constexpr (PAR != 0) heap_sift_up_par<G, PAR == 2>(h, pos, gl);
    else heap_sift_up(h, pos);
  }
  out = item;
  return true;
}

#define EPA_OK 1
#define EPA_NONE 0
#define EPA_OVERFLOW 2

// ccw_face_normal (utils/ccw_face_normal.rs:21-27): None becomes the zero vector.
__device__ __forceinline__ f3 tri_normal(f3 va, f3 vb, f3 vc) {
  f3 nn;
  float nl;
  if (try_normalize(cross(vb - va, vc - va), NCX_FLT_EPS, &nn, &nl)) return nn;
  return f3{0.f, 0.f, 0.f};
}

// A face keeps only its vertex and neighbour ids.  Its normal and the barycentric coordinates
// of the origin's projection are recomputed from the vertices when needed, with the operations
// Face::new used, so they are the same bits as stored ones would be.
__device__ __forceinline__ f3 face_normal(const Arena& A, const FaceD& f) {
  return tri_normal(vpt(A, f.pts[0]), vpt(A, f.pts[1]), vpt(A, f.pts[2]));
}

// Face::new (epa3.rs:96-115) + new_with_proj (:63-94): fills f, returns whether the origin
// projects inside the triangle (proj_is_inside) and the face normal in n.
__device__ __forceinline__ bool make_face(const Arena& A, uint32_t p0, uint32_t p1, uint32_t p2, uint32_t a0,
                                          uint32_t a1, uint32_t a2, FaceD& f, f3& n) {
  f3 va = vpt(A, p0), vb = vpt(A, p1), vc = vpt(A, p2);
  Loc loc;
  proj_triangle(va, vb, vc, loc);
  f.pts[0] = (uint16_t)p0; f.pts[1] = (uint16_t)p1; f.pts[2] = (uint16_t)p2;
  f.adj[0] = (uint16_t)a0; f.adj[1] = (uint16_t)a1; f.adj[2] = (uint16_t)a2;
  f.deleted = 0;
  f.pad = 0;
  n = tri_normal(va, vb, vc);
  return loc.kind == LOC_FACE;
}

// Face::closest_points with the barycentric coordinates new_with_proj stored (zero unless the
// projection was inside the face).
__device__ __forceinline__ void face_points(const Arena& A, const FaceD& f, f3& p1, f3& p2) {
  const CSO a = vget(A, f.pts[0]);
  const CSO b = vget(A, f.pts[1]);
  const CSO c = vget(A, f.pts[2]);
  Loc loc;
  proj_triangle(a.p, b.p, c.p, loc);
  float b0 = 0.f, b1 = 0.f, b2 = 0.f;
  if (loc.kind == LOC_FACE) { b0 = loc.bc0; b1 = loc.bc1; b2 = loc.bc2; }
  p1 = a.o1 * b0 + b.o1 * b1 + c.o1 * b2;
  p2 = a.o2 * b0 + b.o2 * b1 + c.o2 * b2;
}

__device__ __forceinline__ uint32_t next_ccw(const FaceD& f, uint32_t id) {
  if (f.pts[0] == id) return 1;
  if (f.pts[1] == id) return 2;
  return 0;
}

// utils/triangle.rs:85-102
__device__ __forceinline__ bool affinely_dependent(f3 p1, f3 p2, f3 p3) {
  f3 p1p2 = p2 - p1, p1p3 = p3 - p1;
  float e = NCX_FLT_EPS * 100.f;
  float x = norm_sq(cross(p1p2, p1p3));
  float epsilon = e * e;
  if (x == 0.f) return true;
  if (isinf(x)) return false;
  float diff = fabsf(x - 0.f);
  if (diff <= epsilon) return true;
  return diff <= fabsf(x) * NCX_FLT_EPS;
}

// Face::can_be_seen_by (epa3.rs:143-156)
__device__ __forceinline__ bool can_be_seen(const Arena& A, const FaceD& f, uint32_t point, uint32_t opp) {
  f3 v0 = vpt(A, f.pts[0]), v1 = vpt(A, f.pts[1]), v2 = vpt(A, f.pts[2]);
  f3 n = tri_normal(v0, v1, v2);
  f3 p0 = sel3(opp == 0, v0, sel3(opp == 1, v1, v2));
  f3 p1 = sel3(opp == 0, v1, sel3(opp == 1, v2, v0));
  f3 p2 = sel3(opp == 0, v2, sel3(opp == 1, v0, v1));
  f3 pt = vpt(A, point);
  return dot(pt - p0, n) >= -(NCX_FLT_EPS * 10.f) || affinely_dependent(p1, p2, pt);
}

// Suspending and resuming an EPA run between arena tiers.  A run that is about to outgrow its
// arena at an iteration boundary (the vertex capacity reached, or fewer than SUSPEND_MARGIN free
// face slots) saves its polytope -- vertices, faces, heap and loop scalars, bit for bit -- to
// slot i of the next tier's input, where i is its entry in the overflow list, and marks the
// entry with ENTRY_SAVED.  The next tier replays the pair's GJK, which rebuilds the same simplex,
// and continues the EPA from the saved state instead of from that simplex, so only a run that
// outgrows a tier mid-iteration (a silhouette longer than the margin), or finds no free slot,
// starts again.
struct EpaIO {
  const uint8_t* load;  // saved run to continue (null: build the polytope from the simplex)
  uint8_t* save;        // slots for runs this launch suspends (null: it does not suspend)
  size_t save_bytes;    // bytes per slot
  uint32_t save_cap;    // slots
  uint32_t* count;      // the overflow list's counter and entries
  uint32_t* list;
  uint32_t list_cap;
  uint32_t k;           // the pair's index
};
#define EPA_SUSPENDED 3
#define SUSPEND_MARGIN 16u
#define ENTRY_SAVED 0x80000000u
#define SAVE_HEADER_WORDS 8u

__host__ __device__ constexpr size_t epa_save_bytes(uint32_t vcap, uint32_t fcap) {
  return 4 * (size_t)SAVE_HEADER_WORDS + (size_t)vcap * sizeof(VertA) + (size_t)fcap * (sizeof(FaceD) + sizeof(HeapE));
}

// Word copy split over the group's lanes.
template <int G>
__device__ __forceinline__ void copy_words(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, uint32_t n,
                                           int gl) {
  for (uint32_t w = (uint32_t)gl; w < n; w += G) dst[w] = src[w];
}

template <int G>
__device__ __forceinline__ void epa_save(uint8_t* slot, const Arena& A, int gl, uint32_t nv, uint32_t nf, uint32_t hl,
                                         int niter, int iters, HeapE best, float max_dist) {
  uint32_t* d = (uint32_t*)slot;
  if (gl == 0) {
    d[0] = nv; d[1] = nf; d[2] = hl; d[3] = (uint32_t)niter; d[4] = (uint32_t)iters;
    d[5] = __float_as_uint(best.neg_dist); d[6] = best.id; d[7] = __float_as_uint(max_dist);
  }
  d += SAVE_HEADER_WORDS;
  copy_words<G>(d, (const uint32_t*)A.verts, nv * (uint32_t)(sizeof(VertA) / 4), gl);
  d += nv * (sizeof(VertA) / 4);
  copy_words<G>(d, (const uint32_t*)A.faces, nf * (uint32_t)(sizeof(FaceD) / 4), gl);
  d += nf * (sizeof(FaceD) / 4);
  copy_words<G>(d, (const uint32_t*)A.heap, hl * (uint32_t)(sizeof(HeapE) / 4), gl);
}

// Loads a saved run into A (whose capacities are at least the saving tier's); the caller
// orders these writes before the reads of other lanes.
template <int G>
__device__ __forceinline__ void epa_load(const uint8_t* slot, const Arena& A, int gl, uint32_t& nv, uint32_t& nf,
                                         uint32_t& hl, int& niter, int& iters, HeapE& best, float& max_dist) {
  const uint32_t* s = (const uint32_t*)slot;
  nv = s[0]; nf = s[1]; hl = s[2]; niter = (int)s[3]; iters = (int)s[4];
  best = HeapE{__uint_as_float(s[5]), s[6]};
  max_dist = __uint_as_float(s[7]);
  s += SAVE_HEADER_WORDS;
  copy_words<G>((uint32_t*)A.verts, s, nv * (uint32_t)(sizeof(VertA) / 4), gl);
  s += nv * (sizeof(VertA) / 4);
  copy_words<G>((uint32_t*)A.faces, s, nf * (uint32_t)(sizeof(FaceD) / 4), gl);
  s += nf * (sizeof(FaceD) / 4);
  copy_words<G>((uint32_t*)A.heap, s, hl * (uint32_t)(sizeof(HeapE) / 4), gl);
}

// EPA::closest_points (epa3.rs:219-430).  compute_silhouette's recursion
// (:432-454) becomes an explicit stack visiting faces in the same order.  PAR picks how the
// group's lanes share the work: 0 every lane runs the same sequential steps, 1 the lane-parallel
// steps on an LDS arena, 2 the same on a device-memory arena (with workgroup fences).
template <int G, int PAR>
__device__ __forceinline__ int epa(const Ctx2& c, const Simplex& s, const Pool& P, int gl, const Arena& A, f3& p1, f3& p2, f3& n,
                   int& iters, const EpaIO& io) {
  const float eps_tol = NCX_FLT_EPS * 100.f;
  const float gjk_tol = NCX_FLT_EPS * 10.f;
  iters = 0;
  uint32_t nv = 0, nf = 0, hl = 0;
  int niter = 0;
  float max_dist = NCX_FLT_MAX;
  HeapE best;
  if (io.load) {
    epa_load<G>(io.load, A, gl, nv, nf, hl, niter, iters, best, max_dist);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  } else {
    vset(A, nv++, s.v0);
    if (s.dim >= 1) vset(A, nv++, s.v1);
    if (s.dim >= 2) vset(A, nv++, s.v2);
    if (s.dim >= 3) vset(A, nv++, s.v3);
    if (s.dim == 0) {
      p1 = f3{0.f, 0.f, 0.f};
      p2 = f3{0.f, 0.f, 0.f};
      n = f3{0.f, 1.f, 0.f};
      return EPA_OK;
    } else if (s.dim == 3) {
      f3 dp1 = vpt(A, 1) - vpt(A, 0);
      f3 dp2 = vpt(A, 2) - vpt(A, 0);
      f3 dp3 = vpt(A, 3) - vpt(A, 0);
      if (dot(cross(dp1, dp2), dp3) > 0.f) {
        CSO t1 = vget(A, 1), t2 = vget(A, 2);
        vset(A, 1, t2);
        vset(A, 2, t1);
      }
      const uint32_t PT[4][3] = {{0, 1, 2}, {1, 3, 2}, {0, 2, 3}, {0, 3, 1}};
      const uint32_t AD[4][3] = {{3, 1, 2}, {3, 2, 0}, {0, 1, 3}, {2, 1, 0}};
      bool inside[4];
      f3 fn[4];
  #pragma unroll
      for (int f = 0; f < 4; ++f) {
        FaceD fd;
        inside[f] = make_face(A, PT[f][0], PT[f][1], PT[f][2], AD[f][0], AD[f][1], AD[f][2], fd, fn[f]);
        A.faces[nf++] = fd;
      }
  #pragma unroll
      for (int f = 0; f < 4; ++f) {
        if (inside[f]) {
          float dist = dot(fn[f], vpt(A, f));
          if (-dist > gjk_tol) return EPA_NONE;
          A.heap[hl] = HeapE{-dist, (uint32_t)f};
          heap_sift_up(A.heap, hl);
          hl++;
        }
      }
    } else {
      if (s.dim == 1) {
        f3 dpt = vpt(A, 1) - vpt(A, 0);
        // nalgebra orthonormal_subspace_basis for one vector; epa3.rs:299-305 keeps a x dpt.
        f3 a = fabsf(dpt.x) > fabsf(dpt.y) ? f3{dpt.z, 0.f, -dpt.x} : f3{0.f, -dpt.z, dpt.y};
        a = normalize(a);
        vset(A, nv++, cso_from<G>(c, cross(a, dpt), P, gl));
      }
      FaceD f1, f2;
      f3 n1, n2;
      make_face(A, 0, 1, 2, 1, 1, 1, f1, n1);
      make_face(A, 0, 2, 1, 0, 0, 0, f2, n2);
      A.faces[nf++] = f1;
      A.faces[nf++] = f2;
      A.heap[hl] = HeapE{0.f, 0u}; heap_sift_up(A.heap, hl); hl++;
      A.heap[hl] = HeapE{0.f, 1u}; heap_sift_up(A.heap, hl); hl++;
    }
    if (hl == 0) return EPA_NONE;
    best = A.heap[0];
  }
  HeapE fid;
  for (;;) {
    if constexpr (PAR != 0) {
      if (io.save && (nv >= A.vcap || nf + SUSPEND_MARGIN > A.fcap)) {
        uint32_t slot = 0;
        if (gl == 0) slot = atomicAdd(io.count, 1u);
        slot = (uint32_t)__shfl((int)slot, __lane_id() & ~(G - 1), 64);
        const bool saved = slot < io.save_cap;
        if (saved) epa_save<G>(io.save + (size_t)slot * io.save_bytes, A, gl, nv, nf, hl, niter, iters, best, max_dist);
        if (gl == 0 && slot < io.list_cap) io.list[slot] = io.k | (saved ? ENTRY_SAVED : 0u);
        return EPA_SUSPENDED;
      }
    }
    if (!heap_pop<G, PAR>(A.heap, hl, fid, gl)) break;
    FaceD face = A.faces[fid.id];
    if (face.deleted) continue;
    if (nv >= A.vcap) return EPA_OVERFLOW;
    const f3 face_n = face_normal(A, face);
    CSO cso = cso_from<G>(c, face_n, P, gl);
    iters += 1;
    uint32_t sp_id = nv;
    vset(A, nv++, cso);
    float cand = dot(cso.p, face_n);
    if (cand < max_dist) { best = fid; max_dist = cand; }
    float curr_dist = -fid.neg_dist;
    if (max_dist - curr_dist < eps_tol) {
      FaceD bf = A.faces[best.id];
      face_points(A, bf, p1, p2);
      n = face_normal(A, bf);
      return EPA_OK;
    }
    A.faces[fid.id].deleted = 1;
    // compute_silhouette for the three neighbours, depth first.
    uint32_t ns = 0, sp = 0;
    if constexpr (PAR != 0) {
      // Each stack entry's visibility (can_be_seen through the entry's edge) is computed when it
      // is pushed, the entries pushed together by different lanes of the group at once, and kept
      // in bit 15 of the entry: it depends only on the face's vertices, the new point and the
      // edge, none of which change during the search.  The pops, the deleted-face skips and the
      // pushes are the search's own, so the silhouette comes out in the same order.
      const int lane = __lane_id();
      const int gbase = lane & ~(G - 1);
      {
        const int k = 2 - (gl % 3);  // lanes 0, 1, 2 take the entries in push order
        // value selects, not an index into the register copy of the face (that would spill it)
        const uint32_t adj = k == 0 ? face.adj[0] : (k == 1 ? face.adj[1] : face.adj[2]);
        const uint32_t ptk = k == 0 ? face.pts[0] : (k == 1 ? face.pts[1] : face.pts[2]);
        const bool v = can_be_seen(A, A.faces[adj], sp_id, next_ccw(A.faces[adj], ptk));
        const uint64_t vb = __ballot(v) >> gbase;
        for (int j = 0; j < 3; ++j) {
          const uint32_t adjj = face.adj[2 - j];
          const uint32_t oj = next_ccw(A.faces[adjj], face.pts[2 - j]);
          A.stack[sp++] = SilE{(uint16_t)adjj, (uint16_t)(oj | (((vb >> j) & 1ull) ? 0x8000u : 0u))};
        }
      }
      while (sp > 0) {
        const SilE e = A.stack[--sp];
        const uint32_t opp = e.opp & 3u;
        FaceD& fa = A.faces[e.face];
        if (fa.deleted) continue;
        if (!(e.opp & 0x8000u)) {
          if (ns >= A.scap) return EPA_OVERFLOW;
          A.sil[ns++] = SilE{e.face, (uint16_t)opp};
        } else {
          fa.deleted = 1;
          const uint32_t i1 = (opp + 2) % 3, i2 = opp;
          const uint32_t adj1 = fa.adj[i1], adj2 = fa.adj[i2];
          const uint32_t o1 = next_ccw(A.faces[adj1], fa.pts[i1]);
          const uint32_t o2 = next_ccw(A.faces[adj2], fa.pts[i2]);
          if (sp + 2 > A.stcap) return EPA_OVERFLOW;
          const bool odd = (gl & 1) != 0;  // even lanes: adj2 (pushed first), odd lanes: adj1
          const bool v = can_be_seen(A, A.faces[odd ? adj1 : adj2], sp_id, odd ? o1 : o2);
          const uint64_t vb = __ballot(v) >> gbase;
          A.stack[sp++] = SilE{(uint16_t)adj2, (uint16_t)(o2 | ((vb & 1ull) ? 0x8000u : 0u))};
          A.stack[sp++] = SilE{(uint16_t)adj1, (uint16_t)(o1 | ((vb & 2ull) ? 0x8000u : 0u))};
        }
      }
    } else {
      for (int k = 2; k >= 0; --k) {
        uint32_t adj = face.adj[k];
        A.stack[sp++] = SilE{(uint16_t)adj, (uint16_t)next_ccw(A.faces[adj], face.pts[k])};
      }
      while (sp > 0) {
        SilE e = A.stack[--sp];
        FaceD& fa = A.faces[e.face];
        if (fa.deleted) continue;
        if (!can_be_seen(A, fa, sp_id, e.opp)) {
          if (ns >= A.scap) return EPA_OVERFLOW;
          A.sil[ns++] = e;
        } else {
          fa.deleted = 1;
          uint32_t i1 = (e.opp + 2) % 3, i2 = e.opp;
          uint32_t adj1 = fa.adj[i1], adj2 = fa.adj[i2];
          uint32_t o1 = next_ccw(A.faces[adj1], fa.pts[i1]);
          uint32_t o2 = next_ccw(A.faces[adj2], fa.pts[i2]);
          if (sp + 2 > A.stcap) return EPA_OVERFLOW;
          A.stack[sp++] = SilE{(uint16_t)adj2, (uint16_t)o2};
          A.stack[sp++] = SilE{(uint16_t)adj1, (uint16_t)o1};
        }
      }
    }
    uint32_t first_new = nf;
    if (ns == 0) return EPA_NONE;
    if constexpr (PAR != 0) {
      // The new faces are made in parallel, silhouette edge q by lane q mod G of the group, and
      // then checked and pushed in edge order as the loop below does.  Which edges get a face
      // (their silhouette face still live), and so every new face's id, is known before any face
      // is made: making faces only appends faces and rewrites adjacency slots, never the deleted
      // flags or vertex ids read here.  Each edge's results go to its silhouette entry (bit 14:
      // has a face, bit 15: origin projects inside) and to the emptied search stack (the
      // distance).  Lanes of one wavefront see each other's LDS writes in order; a device-memory
      // arena (PAR 2) fences at workgroup scope before the edges' results are read back.
      const int lane = __lane_id();
      const uint64_t gmask = G == 64 ? ~0ull : (((1ull << (G & 63)) - 1ull) << (lane & ~(G - 1)));
      const uint64_t below = (1ull << lane) - 1ull;
      float* sdist = (float*)A.stack;
      uint32_t nvalid = 0;
      for (uint32_t base = 0; base < ns; base += G) {
        const uint32_t q = base + (uint32_t)gl;
        SilE edge = SilE{0, 0};
        bool valid = false;
        if (q < ns) {
          edge = A.sil[q];
          valid = !A.faces[edge.face].deleted;
        }
        const uint64_t vm = __ballot(valid) & gmask;
        const uint32_t new_id = nf + nvalid + (uint32_t)__popcll(vm & below);
        __builtin_amdgcn_wave_barrier();
        if (valid && new_id < A.fcap) {
          const uint32_t pt1 = A.faces[edge.face].pts[(edge.opp + 2) % 3];
          const uint32_t pt2 = A.faces[edge.face].pts[(edge.opp + 1) % 3];
          FaceD nfc;
          f3 nfn;
          const bool inside = make_face(A, pt1, pt2, sp_id, edge.face, new_id + 1, new_id - 1, nfc, nfn);
          A.faces[edge.face].adj[(edge.opp + 1) % 3] = (uint16_t)new_id;
          A.faces[new_id] = nfc;
          sdist[q] = inside ? dot(nfn, vpt(A, pt1)) : 0.f;
          A.sil[q].opp = (uint16_t)(edge.opp | 0x4000u | (inside ? 0x8000u : 0u));
        }
        nvalid += (uint32_t)__popcll(vm);
      }
      __builtin_amdgcn_wave_barrier();
      if constexpr (PAR == 2) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      if (nf + nvalid > A.fcap) return EPA_OVERFLOW;
      for (uint32_t q = 0; q < ns; ++q) {
        const uint32_t fl = A.sil[q].opp;
        if (!(fl & 0x4000u)) continue;
        const uint32_t new_id = nf++;
        if (fl & 0x8000u) {
          const float dist = sdist[q];
          if (dist < curr_dist) {
            face_points(A, face, p1, p2);
            n = face_n;
            return EPA_OK;
          }
          if (-dist > gjk_tol) return EPA_NONE;
          A.heap[hl] = HeapE{-dist, new_id};
          heap_sift_up_par<G, PAR == 2>(A.heap, hl, gl);
          hl++;
        }
      }
    } else {
      for (uint32_t q = 0; q < ns; ++q) {
        SilE edge = A.sil[q];
        if (!A.faces[edge.face].deleted) {
          if (nf >= A.fcap || hl >= A.fcap) return EPA_OVERFLOW;
          uint32_t new_id = nf;
          uint32_t pt1 = A.faces[edge.face].pts[(edge.opp + 2) % 3];
          uint32_t pt2 = A.faces[edge.face].pts[(edge.opp + 1) % 3];
          FaceD nfc;
          f3 nfn;
          bool inside = make_face(A, pt1, pt2, sp_id, edge.face, new_id + 1, new_id - 1, nfc, nfn);
          A.faces[edge.face].adj[(edge.opp + 1) % 3] = (uint16_t)new_id;
          A.faces[nf++] = nfc;
          if (inside) {
            float dist = dot(nfn, vpt(A, nfc.pts[0]));
            if (dist < curr_dist) {
              face_points(A, face, p1, p2);
              n = face_n;
              return EPA_OK;
            }
            if (-dist > gjk_tol) return EPA_NONE;
            A.heap[hl] = HeapE{-dist, new_id};
            heap_sift_up(A.heap, hl);
            hl++;
          }
        }
      }
    }
    if (first_new == nf) return EPA_NONE;
    A.faces[first_new].adj[2] = (uint16_t)(nf - 1);
    A.faces[nf - 1].adj[1] = (uint16_t)first_new;
    niter += 1;
    if (niter > 10000) return EPA_NONE;
  }
  FaceD bf = A.faces[best.id];
  face_points(A, bf, p1, p2);
  n = face_normal(A, bf);
  return EPA_OK;
}

// ------------------------------------------------------------------ per-pair queries
struct PairOut {
  int has;  // 1 = Some(contact); 0 = None; 2 = arena overflow (re-run); 3 = warm-start cap; 4 = suspended (listed)
  f3 w1, w2, n;
  float depth;
  int gjk_iters;
  // The direction of the support-map path's GJKResult (ClosestPoints' or NoIntersection's; the EPA
  // normal; +x when EPA fails), zero when that path did not run or returned nothing
  // (contact_support_map_support_map.rs:62-76): what the polyhedron manifold generator keeps as
  // last_gjk_dir.
  f3 fdir;
};

// contact_support_map_support_map_with_params (contact_support_map_support_map.rs:38-88)
template <int G, int PAR>
__device__ __forceinline__ PairOut contact_sm(const Ctx2& c, float prediction, bool has_init, f3 init_dir, const Pool& P,
                              int gl, const Arena& A, int max_iter, const EpaIO& io) {
  PairOut o;
  o.has = 0;
  o.gjk_iters = 0;
  o.fdir = f3{0.f, 0.f, 0.f};
  f3 dir;
  if (has_init) dir = init_dir;
  else {
    float nl;
    if (!try_normalize(c.m2.t - c.m1.t, NCX_FLT_EPS, &dir, &nl)) dir = f3{1.f, 0.f, 0.f};
  }
  Simplex s;
  sx_init(s);
  sx_reset(s, cso_from<G>(c, dir, P, gl));
  GJKOut r = gjk<G>(c, prediction, s, P, gl, max_iter);
  o.gjk_iters = r.iters;
  if (r.status == GJK_S_CAP && max_iter < 10000) {
    o.has = 3;  // iteration cap of a warm-started run: the caller re-runs plain GJK
    return o;
  }
  f3 w1, w2, n;
  if (r.status == GJK_S_CLOSEST) { w1 = r.p1; w2 = r.p2; n = r.dir; }
  else if (r.status == GJK_S_INTERSECTION) {
    int it;
    int e = epa<G, PAR>(c, s, P, gl, A, w1, w2, n, it, io);
    if (e == EPA_OVERFLOW) { o.has = 2; return o; }
    if (e == EPA_SUSPENDED) { o.has = 4; return o; }
    if (e != EPA_OK) { o.fdir = f3{1.f, 0.f, 0.f}; return o; }  // NoIntersection(x_axis)
  } else {
    if (r.status == GJK_S_NOINTER || r.status == GJK_S_CAP) o.fdir = r.dir;
    return o;
  }
  o.fdir = n;
  o.has = 1;
  o.w1 = w1; o.w2 = w2; o.n = n;
  o.depth = -dot(n, w2 - w1);  // Contact::new_wo_depth (contact.rs:43-46)
  return o;
}

// AABB::local_point_projection + project_point_with_feature for a cuboid (point_aabb.rs:9-135), solid = false.
__device__ __forceinline__ f3 project_cuboid(const iso& m, f3 h, f3 pt, bool& inside_out, int& fkind, int& fid) {
  f3 mins = f3{0.f + (-h.x), 0.f + (-h.y), 0.f + (-h.z)};
  f3 maxs = f3{0.f + h.x, 0.f + h.y, 0.f + h.z};
  f3 ls = inv_xform_pt(m, pt);
  f3 mins_pt = mins - ls, pt_maxs = ls - maxs;
  f3 shift = f3{fmaxf(mins_pt.x, 0.f), fmaxf(mins_pt.y, 0.f), fmaxf(mins_pt.z, 0.f)} -
             f3{fmaxf(pt_maxs.x, 0.f), fmaxf(pt_maxs.y, 0.f), fmaxf(pt_maxs.z, 0.f)};
  bool inside = shift.x == 0.f && shift.y == 0.f && shift.z == 0.f;
  if (inside) {
    float best = -NCX_FLT_MAX;
    bool is_mins = false;
    int best_id = 0;
    for (int i = 0; i < 3; ++i) {
      float a = comp(mins_pt, i), b = comp(pt_maxs, i);
      if (a < b) { if (b > best) { best_id = i; is_mins = false; best = b; } }
      else if (a > best) { best_id = i; is_mins = true; best = a; }
    }
    shift = f3{0.f, 0.f, 0.f};
    set_comp(shift, best_id, is_mins ? best : -best);
  }
  f3 pl = ls + shift;
  inside_out = inside;
  int nzero = 0, last_zero = 0, last_nz = 0;
  for (int i = 0; i < 3; ++i) { if (comp(shift, i) == 0.f) { nzero++; last_zero = i; } else last_nz = i; }
  f3 center = (mins + maxs) * 0.5f;
  if (nzero == 3) {
    fkind = 3; fid = 0;
    for (int i = 0; i < 3; ++i) {
      if (comp(pl, i) > comp(maxs, i) - NCX_FLT_EPS) { fkind = 0; fid = i; break; }
      if (comp(pl, i) <= comp(mins, i) + NCX_FLT_EPS) { fkind = 0; fid = i + 3; break; }
    }
  } else if (nzero == 2) {
    fkind = 0;
    fid = comp(pl, last_nz) < comp(center, last_nz) ? last_nz + 3 : last_nz;
  } else {
    int id = 0;
    for (int i = 0; i < 3; ++i) if (comp(pl, i) < comp(center, i)) id |= 1 << i;
    if (nzero == 0) { fkind = 2; fid = id; } else { fkind = 1; fid = (id << 2) | last_zero; }
  }
  return xform_pt(m, pl);
}

// Cuboid::feature_normal (cuboid.rs:511-561), local frame as in the reference.
__device__ __forceinline__ f3 cuboid_feature_normal(int fkind, int fid) {
  f3 d = f3{0.f, 0.f, 0.f};
  if (fkind == 0) {
    if (fid < 3) set_comp(d, fid, 1.f); else set_comp(d, fid - 3, -1.f);
    return d;
  } else if (fkind == 1) {
    int edge = fid & 3, f1 = (edge + 1) % 3, f2 = (edge + 2) % 3, signs = fid >> 2;
    set_comp(d, f1, (signs & (1 << f1)) ? -1.f : 1.f);
    set_comp(d, f2, (signs & (1 << f2)) ? -1.f : 1.f);
    return normalize(d);
  }
  for (int i = 0; i < 3; ++i) set_comp(d, i, (fid & (1 << i)) ? -1.f : 1.f);
  return normalize(d);
}

// contact_ball_convex_polyhedron (contact_ball_convex_polyhedron.rs:11-58), Cuboid or ConvexHull.
template <int G, int PAR>
__device__ __forceinline__ PairOut contact_ball_convex(f3 c1, float r, const iso& m2, const ShapeD& g2, float prediction,
                                       const Pool& P, int gl, const Arena& A, const EpaIO& io) {
  PairOut o;
  o.has = 0;
  o.gjk_iters = 0;
  o.fdir = f3{0.f, 0.f, 0.f};
  bool inside;
  int fkind = 3, fid = 0;
  f3 world2;
  f3 tri_n = f3{0.f, 1.f, 0.f};  // Triangle::feature_normal: its normal, or +y when degenerate
  if (g2.type == NCX_CUBOID) {
    world2 = project_cuboid(m2, f3{g2.a, g2.b, g2.c}, c1, inside, fkind, fid);
  } else if (g2.type == SHAPE_SEGMENT) {
    // a capsule's segment in the contact-manifold generators: Segment::project_point_with_feature
    world2 = proj_segment_at(m2, f3{0.f, -g2.a, 0.f}, f3{0.f, g2.a, 0.f}, c1, fkind, fid, inside);
  } else if (g2.type == SHAPE_TRIANGLE) {
    // Triangle::project_point_with_feature, solid in 3D (src/query/point/point_triangle.rs:29-51);
    // every location names a feature, so the zero-distance branch below always has a normal
    f3 a, b, c;
    tri_vertices(P, g2.off, a, b, c);
    Loc loc;
    world2 = proj_triangle_at(m2, a, b, c, c1, loc, inside);
    fkind = 0;
    f3 nn;
    float nl;
    if (try_normalize(cross(b - a, c - a), NCX_FLT_EPS, &nn, &nl)) tri_n = nn;  // triangle.rs:137-139, 367-374
  } else {
    // point_projection_on_support_map (point_support_map.rs:15-53), solid = false
    Ctx2 c;
    c.m1 = m2;
    c.m1.t = (-c1) + m2.t;
    c.g1 = g2;
    c.m2 = iso_identity();
    c.g2.type = SHAPE_ORIGIN;
    f3 dir;
    float nl;
    if (!try_normalize(-c.m1.t, NCX_FLT_EPS, &dir, &nl)) dir = f3{1.f, 0.f, 0.f};
    Simplex s;
    sx_init(s);
    sx_reset(s, cso_from<G>(c, dir, P, gl));
    GJKOut r2 = gjk<G>(c, NCX_FLT_MAX, s, P, gl, 10000);
    o.gjk_iters = r2.iters;
    if (r2.status == GJK_S_CLOSEST) {
      inside = false;
      world2 = r2.p1 + c1;
    } else {
      f3 p1, p2, nn;
      int it;
      int e = r2.status == GJK_S_INTERSECTION ? epa<G, PAR>(c, s, P, gl, A, p1, p2, nn, it, io) : EPA_NONE;
      if (e == EPA_OVERFLOW) { o.has = 2; return o; }
      if (e == EPA_SUSPENDED) { o.has = 4; return o; }
      inside = true;
      world2 = e == EPA_OK ? p1 + c1 : c1;
    }
  }
  f3 dpt = world2 - c1;
  f3 dir;
  float dist, depth;
  f3 normal;
  if (try_normalize(dpt, NCX_FLT_EPS, &dir, &dist)) {
    if (inside) { depth = dist + r; normal = -dir; }
    else { depth = -dist + r; normal = dir; }
  } else {
    if (g2.type == SHAPE_TRIANGLE) {
      depth = r;
      normal = -tri_n;
    } else if (g2.type == SHAPE_SEGMENT) {  // a segment location always names a feature
      depth = r;
      normal = -segment_feature_normal(f3{0.f, -g2.a, 0.f}, f3{0.f, g2.a, 0.f}, fkind, fid);
    } else {
      if (g2.type != NCX_CUBOID || fkind == 3) return o;
      depth = r;
      normal = -cuboid_feature_normal(fkind, fid);
    }
  }
  if (depth >= -prediction) {
    o.has = 1;
    o.w1 = c1 + normal * r;
    o.w2 = world2;
    o.n = normal;
    o.depth = depth;
  }
  return o;
}

// query::contact dispatch (contact_shape_shape.rs:10-48).
template <int G, int PAR>
__device__ __forceinline__ PairOut contact_pair(const Ctx2& c, float prediction, bool has_init, f3 init_dir, const Pool& P,
                                int gl, const Arena& A, int max_iter, const EpaIO& io) {
  bool b1 = c.g1.type == NCX_BALL, b2 = c.g2.type == NCX_BALL;
  // ConvexPolyhedron shapes: a ball against one takes the point-projection path (a capsule's
  // segment appears only in the contact-manifold generators, where the ball-polyhedron
  // generator handles it the same way)
  bool poly1 = c.g1.type == NCX_CUBOID || c.g1.type == NCX_CONVEX_HULL || c.g1.type == SHAPE_TRIANGLE ||
               c.g1.type == SHAPE_SEGMENT;
  bool poly2 = c.g2.type == NCX_CUBOID || c.g2.type == NCX_CONVEX_HULL || c.g2.type == SHAPE_TRIANGLE ||
               c.g2.type == SHAPE_SEGMENT;
  if (b1 && b2) {
    // contact_ball_ball (contact_ball_ball.rs:8-38)
    PairOut o;
    o.has = 0;
    o.gjk_iters = 0;
    o.fdir = f3{0.f, 0.f, 0.f};
    float r1 = c.g1.a, r2 = c.g2.a;
    f3 delta = c.m2.t - c.m1.t;
    float d2 = norm_sq(delta);
    float sr = r1 + r2, sre = sr + prediction;
    if (d2 < sre * sre) {
      f3 nrm = d2 != 0.f ? normalize(delta) : f3{1.f, 0.f, 0.f};
      o.has = 1;
      o.w1 = c.m1.t + nrm * r1;
      o.w2 = c.m2.t + nrm * (-r2);
      o.n = nrm;
      o.depth = sr - __builtin_sqrtf(d2);
    }
    return o;
  }
  if (b1 && poly2) return contact_ball_convex<G, PAR>(c.m1.t, c.g1.a, c.m2, c.g2, prediction, P, gl, A, io);
  if (poly1 && b2) {
    PairOut o = contact_ball_convex<G, PAR>(c.m2.t, c.g2.a, c.m1, c.g1, prediction, P, gl, A, io);
    if (o.has == 1) {  // Contact::flip (contact.rs:52-55)
      f3 t = o.w1; o.w1 = o.w2; o.w2 = t;
      o.n = -o.n;
    }
    return o;
  }
  return contact_sm<G, PAR>(c, prediction, has_init, init_dir, P, gl, A, max_iter, io);
}
