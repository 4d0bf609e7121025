// ncx_narrow.hip -- batched convex contact / distance (GJK + EPA) on the GPU.
//
// One group of G lanes (G = 4..64, a power of two inside a 64-wide wavefront)
// handles one pair at a time; the grid loops over pairs.  Every lane of a group
// runs the same GJK / EPA control flow on the same values (results are
// identical across the group), and the lanes split the vertex scans of
// ConvexHull support functions -- the dominant cost -- with an argmax reduction
// that reproduces point_cloud_support_point_id's first-maximum rule
// (src/utils/point_cloud_support_point.rs:6-24).  EPA's growing polytope lives
// in a per-group arena in device memory; a pair whose polytope outgrows the
// arena is re-run by the host in a second launch with a larger arena.
//
// All arithmetic restates the reference in f32 with its operation order:
//   GJK      src/query/algorithms/gjk.rs:76-177
//   simplex  src/query/algorithms/voronoi_simplex3.rs:49-280
//   EPA      src/query/algorithms/epa3.rs:219-454
//   dispatch src/query/contact/contact_shape_shape.rs:10-48,
//            src/query/distance/distance.rs:9-33
#include <algorithm>
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
  if (g2.type == NCX_CUBOID) {
    world2 = project_cuboid(m2, f3{g2.a, g2.b, g2.c}, c1, inside, fkind, fid);
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
    if (g2.type != NCX_CUBOID || fkind == 3) return o;
    depth = r;
    normal = -cuboid_feature_normal(fkind, fid);
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
  bool poly1 = c.g1.type == NCX_CUBOID || c.g1.type == NCX_CONVEX_HULL;
  bool poly2 = c.g2.type == NCX_CUBOID || c.g2.type == NCX_CONVEX_HULL;
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

struct KArgs {
  const ncx_shape* shapes;
  Pool P;
  const ncx_pose* poses;
  const uint32_t* ia;
  const uint32_t* ib;
  const uint32_t* list;  // optional indirection (re-runs)
  size_t npairs;
  float prediction;
  const float* init_dir;
  int max_iter;
  uint8_t* has;
  float* w1;
  float* w2;
  float* nrm;
  float* depth;
  int32_t* iters;
  float* fdir;  // optional: the support-map result direction per pair (PairOut::fdir; the init_dir
                // when that path returned nothing), for a per-pair direction cache
  float* dist;
  uint8_t* arena;
  uint32_t vcap, fcap;
  uint32_t* ovf_count;  // pairs whose EPA outgrew the arena
  uint32_t* ovf_list;   // their pair indices (first ovf_cap of them)
  uint32_t ovf_cap;
  int distance_mode;
  uint32_t nshapes;
  uint32_t* bad_index;  // set to 1 when a pair names a shape >= nshapes
  const uint32_t* n_dev;  // optional device-side pair count (capped by npairs)
  uint32_t* work;         // optional work counter (zero at launch): pairs are taken from it
  const uint32_t* range_b;  // optional device begin and end into list: a run of buckets
  const uint32_t* range_e;
  const uint8_t* sus_load;  // saved EPA runs of the list's ENTRY_SAVED entries, slot = list position
  size_t sus_load_bytes;
  uint8_t* sus_save;        // slots for runs this launch suspends (LDS tiers only; null: none)
  size_t sus_save_bytes;
  uint32_t sus_save_cap;
};

// Contact/distance kernel, BS threads per block in groups of G lanes (one pair per group).
// VC > 0: every group's EPA arena (VC vertices, FC faces, SC silhouette edges, STC search-stack
// entries) lives in LDS; a pair whose polytope outgrows it is listed in ovf for a launch with
// larger arenas.  VC == 0: the arena is the group's slice of a.arena (a.vcap, a.fcap) in device
// memory.
// Arena tiers (capacities; see the launch code below).
#define T1_G 16
#define T1_BS 128
#define T1_VC 44u
#define T1_FC 144u
#define T1_SC 32u
#define T1_STC 32u
#define T2_G 32
#define T2_BS 64
#define T2_BLOCKS_PER_CU (160u * 1024u / ((T2_BS / T2_G) * (uint32_t)arena_bytes(T2_VC, T2_FC, T2_SC, T2_STC)))
#define T2_VC 128u
#define T2_FC 384u
#define T2_SC 64u
#define T2_STC 64u
#define T3_VC 512u
#define T3_FC 1600u
#define T3_SC 128u
#define T3_STC 128u
#define T4_VC 1536u
#define T4_FC 4800u
#define T4_SC 256u
#define T4_STC 256u

constexpr uint32_t ilog2_floor(uint32_t x) { return x <= 1 ? 0u : 1u + ilog2_floor(x >> 1); }

template <int G, int BS, uint32_t VC, uint32_t FC, uint32_t SC, uint32_t STC, int MODE>
__global__ __launch_bounds__(BS, MODE == 1 ? 3 : 2) void k_pairs(KArgs a) {
  const int gl = threadIdx.x & (G - 1);
  const size_t group = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const size_t ngroups = ((size_t)gridDim.x * blockDim.x) / G;
  constexpr bool LDS = VC > 0;
  constexpr size_t LDS_BYTES = LDS ? (BS / G) * arena_bytes(VC, FC, SC, STC) : 16;
  static_assert(LDS_BYTES <= 163840, "LDS arena exceeds the 160 KiB a gfx950 workgroup may hold");
  static_assert(!LDS || (VC < 65536 && FC < 65536 && SC >= 3 && STC >= 3), "arena capacities");
  static_assert(!LDS || ilog2_floor(FC) <= (uint32_t)G, "the parallel heap climb needs a lane per heap level");
  __shared__ __align__(16) uint8_t lds_arena[LDS_BYTES];
  Arena A;
  if constexpr (LDS) A = arena_at(lds_arena, threadIdx.x / G, VC, FC, SC, STC);
  else A = arena_at(a.arena, group, a.vcap, a.fcap, a.fcap, 2 * a.fcap);
  size_t npairs = a.npairs;
  if (a.n_dev) npairs = min(npairs, (size_t)*a.n_dev);
  size_t lbase = 0;
  if (a.range_b) {
    lbase = *a.range_b;
    npairs = *a.range_e - lbase;
  }
  for (size_t it = 0;; ++it) {
    size_t q;
    if (a.work) {  // next pair from the work counter (lane 0 of the group takes it)
      uint32_t t = 0;
      if (gl == 0) t = atomicAdd(a.work, 1u);
      q = (uint32_t)__shfl((int)t, 0, G);
    } else {
      q = group + it * ngroups;
    }
    if (q >= npairs) break;
    const uint32_t entry = a.list ? a.list[lbase + q] : (uint32_t)q;
    size_t k = entry & ~ENTRY_SAVED;
    EpaIO io;
    // Runs are saved by the second to fourth LDS tiers and continued by the third, fourth and
    // device-memory tiers; the first tier, where nearly all pairs finish, has neither path.
    constexpr bool SAVES = LDS && VC > T1_VC;
    constexpr bool LOADS = !LDS || VC > T2_VC;
    io.load = (LOADS && a.list && (entry & ENTRY_SAVED) && a.sus_load) ? a.sus_load + (size_t)q * a.sus_load_bytes
                                                                       : nullptr;
    io.save = SAVES ? a.sus_save : nullptr;
    io.save_bytes = a.sus_save_bytes;
    io.save_cap = a.sus_save_cap;
    io.count = a.ovf_count;
    io.list = a.ovf_list;
    io.list_cap = a.ovf_cap;
    io.k = (uint32_t)k;
    uint32_t i1 = a.ia[k], i2 = a.ib[k];
    if (i1 >= a.nshapes || i2 >= a.nshapes) {
      if (gl == 0) atomicOr(a.bad_index, 1u);
      continue;
    }
    Ctx2 c;
    c.g1 = load_shape(a.shapes, i1);
    c.g2 = load_shape(a.shapes, i2);
    c.m1 = load_pose(a.poses, i1);
    c.m2 = load_pose(a.poses, i2);
    if (MODE == 1) {  // distance (a compile-time branch: each kernel holds one query's code)
      float d = 0.f;
      int it = 0;
      bool ovf = false;
      if (c.g1.type == NCX_BALL && c.g2.type == NCX_BALL) {
        // distance_ball_ball (distance_ball_ball.rs:7-24)
        f3 delta = c.m2.t - c.m1.t;
        float d2 = norm_sq(delta);
        float sr = c.g1.a + c.g2.a;
        d = d2 <= sr * sr ? 0.f : __builtin_sqrtf(d2) - sr;
      } else {
        // distance_support_map_support_map_with_params (distance_support_map_support_map.rs:25-53)
        // init_dir (an all-zero entry means None) with GJK capped at max_iter; on the cap,
        // plain GJK from the default direction judges the pair.
        f3 d0 = c.m1.t - c.m2.t;
        bool warm = false;
        if (a.init_dir) {
          f3 idir = f3{a.init_dir[3 * k], a.init_dir[3 * k + 1], a.init_dir[3 * k + 2]};
          if (idir.x != 0.f || idir.y != 0.f || idir.z != 0.f) {
            d0 = idir;
            warm = true;
          }
        }
        for (int pass = 0; pass < 2; ++pass) {
          f3 dir;
          float nl;
          if (!try_normalize(d0, NCX_FLT_EPS, &dir, &nl)) dir = f3{1.f, 0.f, 0.f};
          Simplex s;
          sx_init(s);
          sx_reset(s, cso_from<G>(c, dir, a.P, gl));
          GJKOut r = gjk<G>(c, NCX_FLT_MAX, s, a.P, gl, warm ? a.max_iter : 10000);
          it += r.iters;
          if (warm && r.status == GJK_S_CAP && a.max_iter < 10000) {
            warm = false;
            d0 = c.m1.t - c.m2.t;
            continue;
          }
          d = r.status == GJK_S_CLOSEST ? norm(r.p1 - r.p2) : 0.f;
          break;
        }
      }
      (void)ovf;
      if (gl == 0) {
        a.dist[k] = d;
        if (a.iters) a.iters[k] = it;
      }
      continue;
    }
    f3 idir = f3{0.f, 0.f, 0.f};
    bool has_init = false;
    if (a.init_dir) {
      idir = f3{a.init_dir[3 * k], a.init_dir[3 * k + 1], a.init_dir[3 * k + 2]};
      // An all-zero entry means None for this pair (the reference's default direction).
      has_init = idir.x != 0.f || idir.y != 0.f || idir.z != 0.f;
    }
    PairOut o = contact_pair<G, LDS ? 1 : 2>(c, a.prediction, has_init, idir, a.P, gl, A, has_init ? a.max_iter : 10000, io);
    int extra_iters = 0;
    if (o.has == 3) {
      // Warm start hit its iteration cap: plain GJK (init_dir = None) judges the pair.
      extra_iters = o.gjk_iters;
      o = contact_pair<G, LDS ? 1 : 2>(c, a.prediction, false, idir, a.P, gl, A, 10000, io);
    }
    if (o.has == 4) continue;  // suspended: listed (and saved) for the next tier
    if (o.has == 2) {
      if (gl == 0) {
        uint32_t slot = atomicAdd(a.ovf_count, 1u);
        if (slot < a.ovf_cap) a.ovf_list[slot] = (uint32_t)k;
      }
      continue;
    }
    if (gl == 0) {
      a.has[k] = (uint8_t)o.has;
      if (o.has == 1) {
        a.w1[3 * k] = o.w1.x; a.w1[3 * k + 1] = o.w1.y; a.w1[3 * k + 2] = o.w1.z;
        a.w2[3 * k] = o.w2.x; a.w2[3 * k + 1] = o.w2.y; a.w2[3 * k + 2] = o.w2.z;
        a.nrm[3 * k] = o.n.x; a.nrm[3 * k + 1] = o.n.y; a.nrm[3 * k + 2] = o.n.z;
        a.depth[k] = o.depth;
      } else {
        a.w1[3 * k] = 0.f; a.w1[3 * k + 1] = 0.f; a.w1[3 * k + 2] = 0.f;
        a.w2[3 * k] = 0.f; a.w2[3 * k + 1] = 0.f; a.w2[3 * k + 2] = 0.f;
        a.nrm[3 * k] = 0.f; a.nrm[3 * k + 1] = 0.f; a.nrm[3 * k + 2] = 0.f;
        a.depth[k] = 0.f;
      }
      if (a.iters) a.iters[k] = o.gjk_iters + extra_iters;
      if (a.fdir) {
        // ConvexPolyhedronConvexPolyhedronManifoldGenerator keeps its last_gjk_dir when the
        // result carries no direction (manifold_generator.rs:115-146)
        const bool z = o.fdir.x == 0.f && o.fdir.y == 0.f && o.fdir.z == 0.f;
        const f3 fd = z ? idir : o.fdir;
        a.fdir[3 * k] = fd.x; a.fdir[3 * k + 1] = fd.y; a.fdir[3 * k + 2] = fd.z;
      }
    }
  }
}

// ------------------------------------------------------------------ host side
__global__ void k_shape_radius(ncx_shape* __restrict__ shapes, size_t n, Pool P);

extern "C" int ncx_shapes_create(ncx_ctx* ctx, const ncx_shape* shapes, size_t n, const float* hx, const float* hy,
                                 const float* hz, size_t npts, ncx_shapes** out) {
  NCX_REQUIRE(ctx && out && (shapes || n == 0), "ncx_shapes_create: null argument");
  for (size_t i = 0; i < n; ++i) {
    const ncx_shape& s = shapes[i];
    NCX_REQUIRE(s.type <= NCX_CONVEX_HULL, "ncx_shapes_create: unknown shape type");
    if (s.type == NCX_CONVEX_HULL)
      NCX_REQUIRE(s.hull_count > 0 && (size_t)s.hull_offset + s.hull_count <= npts,
                  "ncx_shapes_create: hull range outside the point pool");
  }
  ncx_shapes* S = new ncx_shapes;
  S->ctx = ctx;
  S->n = n;
  S->npts = npts;
  int rc;
  if ((rc = S->shapes.reserve(n * sizeof(ncx_shape) + 64)) || (rc = S->hx.reserve(npts * 4 + 4)) ||
      (rc = S->hy.reserve(npts * 4 + 4)) || (rc = S->hz.reserve(npts * 4 + 4)) || (rc = contact_tiers_reserve(S->tiers, 4096))) {
    delete S;
    return rc;
  }
  if (n) NCX_HIP_CHECK(hipMemcpyAsync(S->shapes.p, shapes, n * sizeof(ncx_shape), hipMemcpyHostToDevice, ctx->stream));
  if (npts) {
    NCX_HIP_CHECK(hipMemcpyAsync(S->hx.p, hx, npts * 4, hipMemcpyHostToDevice, ctx->stream));
    NCX_HIP_CHECK(hipMemcpyAsync(S->hy.p, hy, npts * 4, hipMemcpyHostToDevice, ctx->stream));
    NCX_HIP_CHECK(hipMemcpyAsync(S->hz.p, hz, npts * 4, hipMemcpyHostToDevice, ctx->stream));
  }
  if (n) {
    hipLaunchKernelGGL(k_shape_radius, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, S->shapes.as<ncx_shape>(), n,
                       Pool{S->hx.as<float>(), S->hy.as<float>(), S->hz.as<float>()});
    NCX_HIP_CHECK(hipGetLastError());
  }
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  *out = S;
  return NCX_OK;
}

extern "C" int ncx_shapes_destroy(ncx_shapes* s) {
  if (!s) return NCX_OK;
  s->shapes.release(); s->hx.release(); s->hy.release(); s->hz.release(); contact_tiers_release(s->tiers);
  s->qerr.release();
  delete s;
  return NCX_OK;
}

// The autotuner's state for one query (0 contact, 1 distance): per vertex-count bucket the
// chosen variant (0, 1, 2 = 8, 16, 32 lanes per pair; -1 untuned) and each variant's sample time.
extern "C" int ncx_shapes_tuning(ncx_shapes* s, int query, int32_t* choice, float* sample_ms) {
  NCX_REQUIRE(s && choice && (query == 0 || query == 1), "ncx_shapes_tuning: bad argument");
  for (int b = 0; b < NCX_NBUCKETS; ++b) {
    choice[b] = s->tune[query][b];
    if (sample_ms)
      for (int v = 0; v < 3; ++v) sample_ms[3 * b + v] = s->tune_ms[query][b][v];
  }
  return NCX_OK;
}

extern "C" int ncx_shapes_count(ncx_shapes* s, size_t* n) {
  NCX_REQUIRE(s && n, "ncx_shapes_count: null argument");
  *n = s->n;
  return NCX_OK;
}

// ------------------------------------------------------------------ vertex-count buckets
// Sort key of a pair: its vertex-count bucket (by the larger hull vertex count of its shapes: 0 no
// hull, 1 up to 32, 2 up to 96, 3 more; pairs of one bucket have similar support-scan costs and
// get one kernel variant) times 128, plus 8 x its pair of shape types, plus an overlap bin:
// floor(8 x centre distance / (sum of the shapes' bounding radii)), at most 7.  Pairs of one
// type pair take the same dispatch branch, and pairs of one bin are about as likely to need EPA
// (on a 1M-shape scene: bins 0-2 ~100% penetrating, 4: 80%, 6: 31%, 8 and up: none), so the groups
// of a wavefront stay on one code path.  On that scene's 10.2 M pairs contacts took 144 ms in
// their own order, 133 ms sorted by type pair, 121 ms by type pair and bin.  The radius of shape i
// is kept in its record's pad[0] (ncx_shapes_create).
__device__ __forceinline__ uint32_t pair_key(const ncx_shape* __restrict__ shapes, uint32_t nshapes,
                                             const ncx_pose* __restrict__ poses, uint32_t i1, uint32_t i2) {
  if (i1 >= nshapes || i2 >= nshapes) return 0;  // the pair kernel reports the bad index
  const ncx_shape s1 = shapes[i1], s2 = shapes[i2];
  uint32_t c1 = s1.type == NCX_CONVEX_HULL ? s1.hull_count : 0;
  uint32_t c2 = s2.type == NCX_CONVEX_HULL ? s2.hull_count : 0;
  uint32_t m = c1 > c2 ? c1 : c2;
  uint32_t b = m == 0 ? 0u : (m <= 32 ? 1u : (m <= 96 ? 2u : 3u));
  uint32_t t1 = s1.type < 3u ? s1.type : 3u, t2 = s2.type < 3u ? s2.type : 3u;
  const float dx = poses[i1].tx - poses[i2].tx, dy = poses[i1].ty - poses[i2].ty, dz = poses[i1].tz - poses[i2].tz;
  const float rs = __uint_as_float(s1.pad[0]) + __uint_as_float(s2.pad[0]);
  const float bin = (float)NCX_OVERLAP_BINS * __builtin_sqrtf(dx * dx + dy * dy + dz * dz) / rs;
  const uint32_t o = bin < (float)(NCX_OVERLAP_BINS - 1) ? (uint32_t)bin : NCX_OVERLAP_BINS - 1;  // NaN and large: last bin
  return 16u * NCX_OVERLAP_BINS * b + NCX_OVERLAP_BINS * (4u * (t1 < t2 ? t1 : t2) + (t1 < t2 ? t2 : t1)) + o;
}

// Bounding radius of each shape about its local origin, into pad[0] of its device record (a
// sort heuristic only: nothing else reads it).
__global__ __launch_bounds__(256) void k_shape_radius(ncx_shape* __restrict__ shapes, size_t n, Pool P) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ncx_shape s = shapes[i];
  float r = 0.f;
  if (s.type == NCX_BALL) r = s.p0;
  else if (s.type == NCX_CUBOID) r = __builtin_sqrtf(s.p0 * s.p0 + s.p1 * s.p1 + s.p2 * s.p2);
  else if (s.type == NCX_CAPSULE) r = s.p0 + s.p1;
  else {
    for (uint32_t k = 0; k < s.hull_count; ++k) {
      const uint32_t j = s.hull_offset + k;
      r = fmaxf(r, P.x[j] * P.x[j] + P.y[j] * P.y[j] + P.z[j] * P.z[j]);
    }
    r = __builtin_sqrtf(r);
  }
  shapes[i].pad[0] = __float_as_uint(r);
}

#define BKT_BLOCK 256
#define BKT_ITEMS 8

// Pass 1: each pair's key, and per-key counts (LDS histogram per block, one atomic per key per
// block).
__global__ __launch_bounds__(BKT_BLOCK) void k_bucket_count(const uint32_t* __restrict__ ia,
                                                            const uint32_t* __restrict__ ib,
                                                            const ncx_shape* __restrict__ shapes, uint32_t nshapes,
                                                            const ncx_pose* __restrict__ poses, size_t npairs,
                                                            const uint32_t* __restrict__ n_dev,
                                                            uint32_t* __restrict__ hdr, uint32_t* __restrict__ keys) {
  __shared__ uint32_t h[NCX_NKEYS];
  for (uint32_t k = threadIdx.x; k < NCX_NKEYS; k += BKT_BLOCK) h[k] = 0;
  __syncthreads();
  size_t n = npairs;
  if (n_dev) n = min(n, (size_t)*n_dev);
  const size_t base = (size_t)blockIdx.x * BKT_BLOCK * BKT_ITEMS;
  for (int r = 0; r < BKT_ITEMS; ++r) {
    size_t q = base + (size_t)r * BKT_BLOCK + threadIdx.x;
    if (q < n) {
      const uint32_t key = pair_key(shapes, nshapes, poses, ia[q], ib[q]);
      keys[q] = key;
      atomicAdd(&h[key], 1u);
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < NCX_NKEYS; k += BKT_BLOCK)
    if (h[k]) atomicAdd(&hdr[NCX_KEY_CNT + k], h[k]);
}

// Key offsets (the scatter cursors) and each bucket's range [begin, end) in the partitioned list.
__global__ void k_bucket_offsets(uint32_t* __restrict__ hdr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t off = 0;
  for (uint32_t key = 0; key < NCX_NKEYS; ++key) {
    constexpr uint32_t per_bucket = NCX_NKEYS / NCX_NBUCKETS;
    if (key % per_bucket == 0) hdr[NCX_BKT_RANGE + 2 * (key / per_bucket)] = off;
    hdr[NCX_KEY_FILL + key] = off;
    off += hdr[NCX_KEY_CNT + key];
    if (key % per_bucket == per_bucket - 1) hdr[NCX_BKT_RANGE + 2 * (key / per_bucket) + 1] = off;
  }
}

// Pass 2: scatter pair indices to their key's range (one atomic per key per block for the
// block's base, LDS atomics within the block).
__global__ __launch_bounds__(BKT_BLOCK) void k_bucket_scatter(const uint32_t* __restrict__ keys, size_t npairs,
                                                              const uint32_t* __restrict__ n_dev,
                                                              uint32_t* __restrict__ hdr, uint32_t* __restrict__ list) {
  __shared__ uint32_t h[NCX_NKEYS], bb[NCX_NKEYS];
  for (uint32_t k = threadIdx.x; k < NCX_NKEYS; k += BKT_BLOCK) h[k] = 0;
  __syncthreads();
  size_t n = npairs;
  if (n_dev) n = min(n, (size_t)*n_dev);
  const size_t base = (size_t)blockIdx.x * BKT_BLOCK * BKT_ITEMS;
  uint32_t mine[BKT_ITEMS];
#pragma unroll
  for (int r = 0; r < BKT_ITEMS; ++r) {
    size_t q = base + (size_t)r * BKT_BLOCK + threadIdx.x;
    mine[r] = q < n ? keys[q] : 0xFFFFFFFFu;
    if (mine[r] != 0xFFFFFFFFu) atomicAdd(&h[mine[r]], 1u);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < NCX_NKEYS; k += BKT_BLOCK) {
    bb[k] = h[k] ? atomicAdd(&hdr[NCX_KEY_FILL + k], h[k]) : 0u;
    h[k] = 0;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < BKT_ITEMS; ++r) {
    if (mine[r] == 0xFFFFFFFFu) continue;
    size_t q = base + (size_t)r * BKT_BLOCK + threadIdx.x;
    list[bb[mine[r]] + atomicAdd(&h[mine[r]], 1u)] = (uint32_t)q;
  }
}

// Arena tiers, all but the last in LDS and in the graph:
//   T1: 8 pairs per block (8, 16 or 32 lanes per pair, the autotuner's choice), arenas of 44
//       vertices and 144 faces (4.7 KiB per pair; 4 blocks fit a CU);
//   T2: 32 lanes per pair, 2 pairs per 64-thread block, 128 vertices and 384 faces (12.5 KiB;
//       with 16 lanes and 4 pairs per block only 3 blocks fit a CU's LDS, one SIMD idle);
//   T3: 64 lanes, one pair per 64-thread block, 512 vertices and 1600 faces (51 KiB);
//   T4: 64 lanes, one pair per 64-thread block, T4_VC vertices and T4_FC faces (up to the
//       160 KiB a gfx950 workgroup may hold; one block per CU);
//   then device-memory arenas 8x larger than T4's, 64 lanes per pair, host-driven.

// Saved-run slots handed from each LDS tier to the next (the first SUS_CAP[t] suspended runs of
// a launch are saved; later ones start again in the next tier) and their sizes.  The first tier
// saves nothing (slot count 0): replaying its short runs costs less than carrying the save path
// in its kernel.
static const uint32_t SUS_CAP[4] = {0, 4096, 1024, 256};
static const size_t SUS_BYTES[4] = {epa_save_bytes(T1_VC, T1_FC), epa_save_bytes(T2_VC, T2_FC),
                                    epa_save_bytes(T3_VC, T3_FC), epa_save_bytes(T4_VC, T4_FC)};

// Overflow lists of at least cap entries each (at most 2^24: a call with more overflowing pairs
// fails with NCX_E_CAPACITY), and the saved-run slots.
int contact_tiers_reserve(ContactTiers& t, size_t cap) {
  for (int i = 0; i < 4; ++i) {
    if (t.sus[i].p || SUS_CAP[i] == 0) continue;
    int rc = t.sus[i].reserve((size_t)SUS_CAP[i] * SUS_BYTES[i]);
    if (rc) return rc;
  }
  cap = std::min<size_t>(std::max<size_t>(cap, 4096), (size_t)1 << 24);
  if (cap <= t.cap) return NCX_OK;
  t.ovf.release();
  t.cap = 0;
  int rc = t.ovf.reserve(4 * (NCX_OVF_HEADER + 4 * cap));
  if (rc) return rc;
  t.cap = (uint32_t)cap;
  return NCX_OK;
}

void contact_tiers_release(ContactTiers& t) {
  t.ovf.release();
  t.blist.release();
  t.bkey.release();
  t.bcap = 0;
  t.cap = 0;
  for (int i = 0; i < 4; ++i) t.sus[i].release();
  delete t.last;
  t.last = nullptr;
}

// Device-memory tiers (VC = 0): static grid-stride over the listed pairs.
static int launch_global_tier(hipStream_t stream, KArgs a, size_t npairs) {
  a.npairs = npairs;
  a.work = nullptr;
  hipLaunchKernelGGL((k_pairs<64, 128, 0u, 0u, 0u, 0u, 0>), dim3(ceil_div(npairs * 64, 128)), dim3(128), 0, stream, a);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

// Re-runs the n pairs listed at list_src (device) with device-memory arenas 8x larger per round
// (starting from 8 x the last tier's vcap/fcap) until none overflows; host-driven (synchronises
// the stream).  Chunks keep each round's arenas <= 2 GiB.  The first round continues the runs
// the last LDS tier saved (sus_src, slot = list position); device-memory rounds do not suspend.
static int rerun_overflows(ContactTiers& T, ncx_ctx* ctx, KArgs a, uint32_t n, const uint32_t* list_src,
                           uint32_t vcap, uint32_t fcap, const uint8_t* sus_src, size_t sus_bytes) {
  if (n == 0) return NCX_OK;
  uint32_t* ovf = T.ovf.as<uint32_t>();
  DevBuf lst;
  int rc;
  if ((rc = lst.reserve((size_t)n * 4))) return rc;
  NCX_HIP_CHECK(hipMemcpyAsync(lst.p, list_src, (size_t)n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  while (n > 0) {
    vcap *= 8;
    fcap *= 8;
    if (fcap > 65535) fcap = 65535;
    if (vcap > 32767) vcap = 32767;
    const size_t per = arena_bytes(vcap, fcap);
    const size_t chunk = std::max<size_t>(1, ((size_t)2 << 30) / per);
    DevBuf big;
    if ((rc = big.reserve(std::min<size_t>(n, chunk) * per))) return rc;
    if ((rc = fill_u32(ctx, ovf + NCX_OVF_T2_COUNT, 1, 0u))) return rc;
    for (size_t off = 0; off < n; off += chunk) {
      size_t m = std::min<size_t>(chunk, n - off);
      KArgs b = a;
      b.list = lst.as<uint32_t>() + off;
      b.n_dev = nullptr;
      b.arena = big.as<uint8_t>();
      b.vcap = vcap;
      b.fcap = fcap;
      b.ovf_count = ovf + NCX_OVF_T2_COUNT;
      b.ovf_list = T.t2_list();
      b.ovf_cap = T.cap;
      b.sus_load = sus_src ? sus_src + off * sus_bytes : nullptr;
      b.sus_load_bytes = sus_bytes;
      b.sus_save = nullptr;
      b.sus_save_bytes = 0;
      b.sus_save_cap = 0;
      if ((rc = launch_global_tier(ctx->stream, b, m))) return rc;
    }
    sus_src = nullptr;  // later rounds list this tier's own overflows, with nothing saved
    uint32_t left = 0;
    NCX_HIP_CHECK(hipMemcpyAsync(&left, ovf + NCX_OVF_T2_COUNT, 4, hipMemcpyDeviceToHost, ctx->stream));
    NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    big.release();
    if (left) NCX_REQUIRE(fcap < 65535, "contact batch: EPA polytope exceeds 65535 faces");
    n = left;
    if (n) {
      NCX_REQUIRE(n <= T.cap, "contact batch: too many EPA arena overflows");
      NCX_HIP_CHECK(hipMemcpyAsync(lst.p, T.t2_list(), (size_t)n * 4, hipMemcpyDeviceToDevice,
                                   ctx->stream));
    }
  }
  return NCX_OK;
}

// First-tier kernel variants, the autotuner's choices: group width G = 8, 16, 32 lanes per
// pair.  Contact variants all hold 8 pairs per block (64, 128, 256 threads) with the same
// 44-vertex, 144-face LDS arena, so a variant changes occupancy and lanes per pair but not which
// pairs overflow to the second tier.  Distance queries need no EPA arena and use 128-thread
// variants without LDS.
static const int T1_VARIANT_G[3] = {8, 16, 32};
static const int T1_VARIANT_BS[2][3] = {{64, 128, 256}, {128, 128, 128}};
static const unsigned T1_BLOCKS_PER_CU[2][3] = {{8, 8, 4}, {8, 8, 8}};  // twice the resident blocks

static int launch_t1(int variant, bool distance, hipStream_t stream, const KArgs& a, unsigned blocks) {
  if (distance) {
    switch (variant) {
      case 0: hipLaunchKernelGGL((k_pairs<8, 128, 0u, 0u, 0u, 0u, 1>), dim3(blocks), dim3(128), 0, stream, a); break;
      case 2: hipLaunchKernelGGL((k_pairs<32, 128, 0u, 0u, 0u, 0u, 1>), dim3(blocks), dim3(128), 0, stream, a); break;
      default: hipLaunchKernelGGL((k_pairs<16, 128, 0u, 0u, 0u, 0u, 1>), dim3(blocks), dim3(128), 0, stream, a); break;
    }
  } else {
    switch (variant) {
      case 0:
        hipLaunchKernelGGL((k_pairs<8, 64, T1_VC, T1_FC, T1_SC, T1_STC, 0>), dim3(blocks), dim3(64), 0, stream, a);
        break;
      case 2:
        hipLaunchKernelGGL((k_pairs<32, 256, T1_VC, T1_FC, T1_SC, T1_STC, 0>), dim3(blocks), dim3(256), 0, stream, a);
        break;
      default:
        hipLaunchKernelGGL((k_pairs<16, 128, T1_VC, T1_FC, T1_SC, T1_STC, 0>), dim3(blocks), dim3(128), 0, stream, a);
        break;
    }
  }
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

// Saved-run buffers of LDS tier `tier` (0 = T1): it continues the runs T(tier) saved and saves
// its own suspended runs for the next tier.
static void set_sus(ContactTiers& T, KArgs& b, int tier) {
  b.sus_load = tier > 0 ? T.sus[tier - 1].as<uint8_t>() : nullptr;
  b.sus_load_bytes = tier > 0 ? SUS_BYTES[tier - 1] : 0;
  b.sus_save = T.sus[tier].as<uint8_t>();
  b.sus_save_bytes = SUS_BYTES[tier];
  b.sus_save_cap = T.sus[tier].p ? SUS_CAP[tier] : 0u;
}

// T2 over the device-counted T1 overflow list.
static KArgs t2_args(ContactTiers& T, const KArgs& a) {
  uint32_t* ovf = T.ovf.as<uint32_t>();
  KArgs b = a;
  b.range_b = nullptr;
  b.range_e = nullptr;
  b.list = T.t1_list();
  b.n_dev = ovf + NCX_OVF_T1_COUNT;
  b.npairs = T.cap;
  b.vcap = T2_VC;
  b.fcap = T2_FC;
  b.ovf_count = ovf + NCX_OVF_T2_COUNT;
  b.ovf_list = T.t2_list();
  b.work = ncx_work_set(ovf, NCX_WORK_T2);
  set_sus(T, b, 1);
  return b;
}

// T3 over the device-counted T2 overflow list: 64 lanes and one pair per block, with a
// 512-vertex, 1600-face LDS arena (51 KiB).
static KArgs t3_args(ContactTiers& T, const KArgs& a) {
  uint32_t* ovf = T.ovf.as<uint32_t>();
  KArgs b = t2_args(T, a);
  b.list = T.t2_list();
  b.n_dev = ovf + NCX_OVF_T2_COUNT;
  b.vcap = T3_VC;
  b.fcap = T3_FC;
  b.ovf_count = ovf + NCX_OVF_T3_COUNT;
  b.ovf_list = T.t3_list();
  b.work = ncx_work_set(ovf, NCX_WORK_T3);
  set_sus(T, b, 2);
  return b;
}

// T4 over the device-counted T3 overflow list: one pair per 64-thread block with the largest
// LDS arena a workgroup may hold.
static KArgs t4_args(ContactTiers& T, const KArgs& a) {
  uint32_t* ovf = T.ovf.as<uint32_t>();
  KArgs b = t3_args(T, a);
  b.list = T.t3_list();
  b.n_dev = ovf + NCX_OVF_T3_COUNT;
  b.vcap = T4_VC;
  b.fcap = T4_FC;
  b.ovf_count = ovf + NCX_OVF_T4_COUNT;
  b.ovf_list = T.t4_list();
  b.work = ncx_work_set(ovf, NCX_WORK_T4);
  set_sus(T, b, 3);
  return b;
}

static int launch_t2(ncx_ctx* ctx, ContactTiers& T, const KArgs& a) {
  KArgs b = t2_args(T, a);
  hipLaunchKernelGGL((k_pairs<T2_G, T2_BS, T2_VC, T2_FC, T2_SC, T2_STC, 0>), dim3(ctx->num_cus * T2_BLOCKS_PER_CU), dim3(T2_BS), 0,
                     ctx->stream, b);
  NCX_HIP_CHECK(hipGetLastError());
  KArgs c = t3_args(T, a);
  hipLaunchKernelGGL((k_pairs<64, 64, T3_VC, T3_FC, T3_SC, T3_STC, 0>), dim3(ctx->num_cus * 3), dim3(64), 0,
                     ctx->stream, c);
  NCX_HIP_CHECK(hipGetLastError());
  KArgs d = t4_args(T, a);
  hipLaunchKernelGGL((k_pairs<64, 64, T4_VC, T4_FC, T4_SC, T4_STC, 0>), dim3(ctx->num_cus), dim3(64), 0,
                     ctx->stream, d);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

int contact_tiers_reserve_list(ContactTiers& t, size_t npairs) {
  if (npairs <= t.bcap) return NCX_OK;
  t.blist.release();
  t.bkey.release();
  t.bcap = 0;
  int rc = t.blist.reserve(npairs * 4 + 64);
  if (!rc) rc = t.bkey.reserve(npairs * 4 + 64);
  if (rc) return rc;
  t.bcap = npairs;
  return NCX_OK;
}

// Autotuner: for each bucket with enough pairs and no choice yet, time every first-tier variant
// on a sample of the bucket's pairs (HIP events; the least of three rounds after a warm-up
// round, the variants interleaved within each round) and keep the fastest.  The
// sample's results are real results (the full launch recomputes them identically).  Host-driven;
// the choice is kept in the shape set for every later call and scene step.
// Clears the tiers' counts and work counters (not the buckets').
static int clear_tier_counters(ncx_ctx* ctx, uint32_t* hdr) {
  int rc = fill_u32(ctx, hdr, NCX_OVF_TIER_WORDS, 0u);
  if (rc) return rc;
  static_assert(NCX_WORK_TUNE == 0 && NCX_WORK_T2 == 1 && NCX_WORK_T3 == 2 && NCX_WORK_T4 == 3, "work sets");
  return fill_u32(ctx, ncx_work_set(hdr, 0), 4 * NCX_WORK_SET, 0u);
}

#define TUNE_MIN_PAIRS 4096u
#define PARTITION_MIN_PAIRS 16384u
#define TUNE_SAMPLE 131072u
#define TUNE_ROUNDS 4
// Every stride-th entry of a bucket's run of the partitioned list, so the sample holds each of
// the bucket's type pairs in proportion.
__global__ void k_tune_sample(const uint32_t* __restrict__ list, uint32_t begin, uint32_t stride, uint32_t n,
                              uint32_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = list[begin + (size_t)i * stride];
}

static int autotune(ncx_shapes* S, ContactTiers& T, ncx_ctx* ctx, const KArgs& base, int mode) {
  uint32_t* hdr = T.ovf.as<uint32_t>();
  uint32_t rng[2 * NCX_NBUCKETS];
  NCX_HIP_CHECK(hipMemcpyAsync(rng, hdr + NCX_BKT_RANGE, sizeof(rng), hipMemcpyDeviceToHost, ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  DevBuf dr, smp;
  int rc;
  if ((rc = dr.reserve(16)) || (rc = smp.reserve((size_t)TUNE_SAMPLE * 4))) return rc;
  hipEvent_t e0, e1;
  NCX_HIP_CHECK(hipEventCreate(&e0));
  NCX_HIP_CHECK(hipEventCreate(&e1));
  for (int b = 0; b < NCX_NBUCKETS; ++b) {
    const uint32_t cnt = rng[2 * b + 1] - rng[2 * b];
    if (S->tune[mode][b] >= 0 || cnt < TUNE_MIN_PAIRS) continue;
    const uint32_t ns = std::min(cnt, TUNE_SAMPLE);
    hipLaunchKernelGGL(k_tune_sample, dim3(ceil_div(ns, 256)), dim3(256), 0, ctx->stream, base.list, rng[2 * b],
                       cnt / ns, ns, smp.as<uint32_t>());
    NCX_HIP_CHECK(hipGetLastError());
    const uint32_t sr[2] = {0, ns};
    NCX_HIP_CHECK(hipMemcpyAsync(dr.p, sr, 8, hipMemcpyHostToDevice, ctx->stream));
    // Rounds over the three variants, so clock ramps and noise fall on every variant alike; the
    // first round warms up, each variant keeps its least time of the others.
    float ms_v[3] = {0.f, 0.f, 0.f};
    for (int rep = 0; rep < TUNE_ROUNDS && !rc; ++rep) {
      for (int v = 0; v < 3; ++v) {
        KArgs a = base;
        a.list = smp.as<uint32_t>();
        a.range_b = dr.as<uint32_t>();
        a.range_e = dr.as<uint32_t>() + 1;
        a.work = ncx_work_set(hdr, NCX_WORK_TUNE);
        if ((rc = clear_tier_counters(ctx, hdr))) break;
        NCX_HIP_CHECK(hipEventRecord(e0, ctx->stream));
        if ((rc = launch_t1(v, mode == 1, ctx->stream, a, (unsigned)ctx->num_cus * T1_BLOCKS_PER_CU[mode][v]))) break;
        // the variant's first-tier overflows cost a second-tier run: time it too
        if (mode == 0 && (rc = launch_t2(ctx, T, a))) break;
        NCX_HIP_CHECK(hipEventRecord(e1, ctx->stream));
        NCX_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        NCX_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (rep == 1 || (rep > 1 && ms < ms_v[v])) ms_v[v] = ms;
      }
    }
    int best = 1;
    for (int v = 0; v < 3 && !rc; ++v) {
      S->tune_ms[mode][b][v] = ms_v[v];
      if (ms_v[v] < ms_v[best]) best = v;
    }
    if (rc) break;
    S->tune[mode][b] = (int8_t)best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dr.release();
  smp.release();
  if (rc) return rc;
  // the samples left entries in the overflow lists and moved the work counter: start clean
  return clear_tier_counters(ctx, hdr);
}

// The launch: partition the pairs into vertex-count buckets (device-side counts, so it is
// graph-safe), run each bucket with its variant (T1, LDS arenas), then T2, T3 and T4
// rerun the device-counted overflow list of the tier before with larger LDS arenas, the last
// listing what outgrows it for contacts_resolve_overflow.  Every launch takes pairs from a work counter, so uneven pair
// costs balance across the chip.  npairs caps the pair count; n_dev (optional) holds the actual
// count on the device.  may_tune: run the autotuner first if a bucket has no choice yet (host
// synchronisation; never inside graph capture).
static int launch_tiers(ncx_shapes* S, ContactTiers& T, ncx_ctx* ctx, KArgs a, size_t npairs,
                        const uint32_t* n_dev, bool may_tune) {
  uint32_t* ovf = T.ovf.as<uint32_t>();
  NCX_REQUIRE(ovf, "contact launch: overflow lists not allocated");
  NCX_REQUIRE(npairs < ((size_t)1 << 31), "contact launch: more than 2^31 pairs");
  int rc;
  if ((rc = fill_u32(ctx, ovf, NCX_OVF_HEADER, 0u))) return rc;
  const int mode = a.distance_mode ? 1 : 0;
  a.nshapes = (uint32_t)S->n;
  a.bad_index = ovf + NCX_OVF_BAD;
  a.arena = nullptr;
  a.vcap = T1_VC;
  a.fcap = T1_FC;
  a.ovf_count = ovf + NCX_OVF_T1_COUNT;
  a.ovf_list = T.t1_list();
  a.ovf_cap = T.cap;
  set_sus(T, a, 0);
  a.range_b = nullptr;
  a.range_e = nullptr;
  bool untuned = false;
  for (int b = 0; b < NCX_NBUCKETS; ++b) untuned = untuned || S->tune[mode][b] < 0;
  const bool tune_now = may_tune && untuned;
  bool same = true;  // one variant for every bucket
  for (int b = 1; b < NCX_NBUCKETS; ++b) same = same && S->tune[mode][b] == S->tune[mode][0];
  // Partition (sort by key) unless the batch is small and needs no per-bucket variants.
  const bool partition = tune_now || !same || npairs >= PARTITION_MIN_PAIRS;
  if (partition) {
    // partition the pair indices by sort key (vertex-count bucket, then type pair)
    if ((rc = contact_tiers_reserve_list(T, npairs))) return rc;
    unsigned bblocks = std::max(1u, ceil_div(npairs, (size_t)BKT_BLOCK * BKT_ITEMS));
    hipLaunchKernelGGL(k_bucket_count, dim3(bblocks), dim3(BKT_BLOCK), 0, ctx->stream, a.ia, a.ib, a.shapes,
                       a.nshapes, a.poses, npairs, n_dev, ovf, T.bkey.as<uint32_t>());
    NCX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_bucket_offsets, dim3(1), dim3(1), 0, ctx->stream, ovf);
    NCX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_bucket_scatter, dim3(bblocks), dim3(BKT_BLOCK), 0, ctx->stream, T.bkey.as<uint32_t>(),
                       npairs, n_dev, ovf, T.blist.as<uint32_t>());
    NCX_HIP_CHECK(hipGetLastError());
    KArgs t = a;
    t.list = T.blist.as<uint32_t>();
    t.n_dev = nullptr;
    t.npairs = npairs;
    if (tune_now && (rc = autotune(S, T, ctx, t, mode))) return rc;
  }
  if (!partition) {
    // one launch over all pairs, in their own order
    const int v = S->tune[mode][0] >= 0 ? S->tune[mode][0] : 1;
    KArgs ab = a;
    ab.list = nullptr;
    ab.n_dev = n_dev;
    ab.npairs = npairs;
    ab.work = ncx_work_set(ovf, NCX_WORK_BKT);
    unsigned blocks = (unsigned)ctx->num_cus * T1_BLOCKS_PER_CU[mode][v];
    if (!n_dev)
      blocks = std::max(1u, std::min<unsigned>(blocks, ceil_div(npairs * T1_VARIANT_G[v], T1_VARIANT_BS[mode][v])));
    if ((rc = launch_t1(v, mode == 1, ctx->stream, ab, blocks))) return rc;
  } else {
    // one launch per run of consecutive buckets sharing a variant, over the partitioned list
    for (int b0 = 0; b0 < NCX_NBUCKETS;) {
      const int v = S->tune[mode][b0] >= 0 ? S->tune[mode][b0] : 1;
      int b1 = b0;
      while (b1 + 1 < NCX_NBUCKETS && (S->tune[mode][b1 + 1] >= 0 ? S->tune[mode][b1 + 1] : 1) == v) ++b1;
      KArgs ab = a;
      ab.list = T.blist.as<uint32_t>();
      ab.n_dev = nullptr;
      ab.npairs = npairs;
      ab.range_b = ovf + NCX_BKT_RANGE + 2 * b0;
      ab.range_e = ovf + NCX_BKT_RANGE + 2 * b1 + 1;
      ab.work = ncx_work_set(ovf, NCX_WORK_BKT + b0);
      unsigned blocks = (unsigned)ctx->num_cus * T1_BLOCKS_PER_CU[mode][v];
      if (!n_dev)
        blocks = std::max(1u, std::min<unsigned>(blocks, ceil_div(npairs * T1_VARIANT_G[v], T1_VARIANT_BS[mode][v])));
      if ((rc = launch_t1(v, mode == 1, ctx->stream, ab, blocks))) return rc;
      b0 = b1 + 1;
    }
  }
  if (mode == 0 && (rc = launch_t2(ctx, T, a))) return rc;
  if (!T.last) T.last = new KArgs;
  *T.last = t4_args(T, a);
  return NCX_OK;
}

// Graph-safe tiered contact launch over a device-counted pair list (the scene step); the caller
// runs contacts_resolve_overflow with the same tiers after the work completes.
int contacts_launch_device_count(ncx_shapes* S, ContactTiers& T, ncx_ctx* ctx, const ncx_pose* poses,
                                 const uint32_t* ia, const uint32_t* ib, const uint32_t* n_dev, size_t cap,
                                 float prediction, const float* init_dir, int max_iter, uint8_t* has, float* w1,
                                 float* w2, float* nrm, float* depth, int32_t* iters, bool may_tune, float* fdir) {
  KArgs a = {};
  a.fdir = fdir;
  a.shapes = S->shapes.as<ncx_shape>();
  a.P = Pool{S->hx.as<float>(), S->hy.as<float>(), S->hz.as<float>()};
  a.poses = poses;
  a.ia = ia;
  a.ib = ib;
  a.prediction = prediction;
  a.init_dir = init_dir;
  a.max_iter = max_iter;
  a.has = has; a.w1 = w1; a.w2 = w2; a.nrm = nrm; a.depth = depth; a.iters = iters;
  a.distance_mode = 0;
  return launch_tiers(S, T, ctx, a, cap, n_dev, may_tune);
}

// After the last launch_tiers on T has completed on ctx's stream: fails on a bad pair index or on
// more overflows than the lists hold, re-runs what T4 could not hold with device-memory arenas,
// and reports how many pairs needed more than the T1 arena.
int contacts_resolve_overflow(ContactTiers& T, ncx_ctx* ctx, size_t* n_large) {
  NCX_REQUIRE(T.last, "contact resolve: no tiered launch to resolve");
  uint32_t hdr[NCX_OVF_COUNTS];
  NCX_HIP_CHECK(hipMemcpyAsync(hdr, T.ovf.p, sizeof(hdr), hipMemcpyDeviceToHost, ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  NCX_REQUIRE(hdr[NCX_OVF_BAD] == 0, "contact batch: a pair index is not a shape of the set");
  if (hdr[NCX_OVF_T1_COUNT] > T.cap || hdr[NCX_OVF_T2_COUNT] > T.cap || hdr[NCX_OVF_T3_COUNT] > T.cap ||
      hdr[NCX_OVF_T4_COUNT] > T.cap) {
    ncx_set_error("contact batch: more EPA arena overflows in one call than the overflow list holds");
    return NCX_E_CAPACITY;
  }
  if (n_large) *n_large = hdr[NCX_OVF_T1_COUNT];
  return rerun_overflows(T, ctx, *T.last, hdr[NCX_OVF_T4_COUNT], T.t4_list(), T4_VC, T4_FC, T.sus[3].as<uint8_t>(),
                         SUS_BYTES[3]);
}

static int run_pairs(ncx_shapes* S, KArgs a, size_t npairs) {
  if (npairs == 0) return NCX_OK;
  int rc;
  if ((rc = contact_tiers_reserve(S->tiers, npairs))) return rc;
  if ((rc = launch_tiers(S, S->tiers, S->ctx, a, npairs, nullptr, true))) return rc;
  return contacts_resolve_overflow(S->tiers, S->ctx, nullptr);
}

extern "C" int ncx_contact_batch(ncx_shapes* s, const ncx_pose* poses, const uint32_t* ia, const uint32_t* ib,
                                 size_t npairs, float prediction, const float* init_dir, uint8_t* has, float* w1,
                                 float* w2, float* nrm, float* depth, int32_t* iters) {
  NCX_REQUIRE(s && poses && ia && ib && has && w1 && w2 && nrm && depth, "ncx_contact_batch: null argument");
  KArgs a = {};
  a.shapes = s->shapes.as<ncx_shape>();
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>()};
  a.poses = poses;
  a.ia = ia;
  a.ib = ib;
  a.prediction = prediction;
  a.init_dir = init_dir;
  a.max_iter = 10000;
  a.has = has;
  a.w1 = w1;
  a.w2 = w2;
  a.nrm = nrm;
  a.depth = depth;
  a.iters = iters;
  a.distance_mode = 0;
  return run_pairs(s, a, npairs);
}

// Warm-started variant used by the learned initial direction: GJK from init_dir is capped at
// max_iter iterations, after which plain GJK (init_dir = None) recomputes the pair.
extern "C" int ncx_contact_batch_capped(ncx_shapes* s, const ncx_pose* poses, const uint32_t* ia, const uint32_t* ib,
                                        size_t npairs, float prediction, const float* init_dir, int32_t max_iter,
                                        uint8_t* has, float* w1, float* w2, float* nrm, float* depth,
                                        int32_t* iters) {
  NCX_REQUIRE(s && poses && ia && ib && has && w1 && w2 && nrm && depth, "ncx_contact_batch_capped: null argument");
  NCX_REQUIRE(max_iter >= 1 && max_iter <= 10000, "ncx_contact_batch_capped: max_iter out of range");
  KArgs a = {};
  a.shapes = s->shapes.as<ncx_shape>();
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>()};
  a.poses = poses;
  a.ia = ia;
  a.ib = ib;
  a.prediction = prediction;
  a.init_dir = init_dir;
  a.max_iter = init_dir ? max_iter : 10000;
  a.has = has;
  a.w1 = w1;
  a.w2 = w2;
  a.nrm = nrm;
  a.depth = depth;
  a.iters = iters;
  a.distance_mode = 0;
  return run_pairs(s, a, npairs);
}

extern "C" int ncx_distance_batch(ncx_shapes* s, const ncx_pose* poses, const uint32_t* ia, const uint32_t* ib,
                                  size_t npairs, float* dist, int32_t* iters) {
  NCX_REQUIRE(s && poses && ia && ib && dist, "ncx_distance_batch: null argument");
  KArgs a = {};
  a.shapes = s->shapes.as<ncx_shape>();
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>()};
  a.poses = poses;
  a.ia = ia;
  a.ib = ib;
  a.max_iter = 10000;
  a.dist = dist;
  a.iters = iters;
  a.distance_mode = 1;
  return run_pairs(s, a, npairs);
}

// distance_support_map_support_map_with_params with init_dir: GJK from init_dir capped at
// max_iter iterations, after which plain GJK (init_dir = None) recomputes the pair.
extern "C" int ncx_distance_batch_capped(ncx_shapes* s, const ncx_pose* poses, const uint32_t* ia,
                                         const uint32_t* ib, size_t npairs, const float* init_dir, int32_t max_iter,
                                         float* dist, int32_t* iters) {
  NCX_REQUIRE(s && poses && ia && ib && dist, "ncx_distance_batch_capped: null argument");
  NCX_REQUIRE(max_iter >= 1 && max_iter <= 10000, "ncx_distance_batch_capped: max_iter out of range");
  KArgs a = {};
  a.shapes = s->shapes.as<ncx_shape>();
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>()};
  a.poses = poses;
  a.ia = ia;
  a.ib = ib;
  a.init_dir = init_dir;
  a.max_iter = init_dir ? max_iter : 10000;
  a.dist = dist;
  a.iters = iters;
  a.distance_mode = 1;
  return run_pairs(s, a, npairs);
}

// ------------------------------------------------------------------ shape AABBs
// One thread per shape, except hulls: a wavefront takes its hull lanes four at a time, 16 lanes
// per hull, each lane transforming four of the hull's points per pass with independent loads
// (coalesced across the lanes: the point pool is SoA), and reduces min/max over its 16 lanes.
// A thread looping over its own hull paid one memory round trip per point; one hull per
// wavefront at a time paid one per hull.  The min/max of a set does not depend on the visiting
// order (a point read twice, to keep the loads unconditional, changes nothing).
__device__ __forceinline__ float grp16_min(float x) {
#pragma unroll
  for (int s = 8; s >= 1; s >>= 1) x = fminf(x, __shfl_xor(x, s, 64));
  return x;
}
__device__ __forceinline__ float grp16_max(float x) {
#pragma unroll
  for (int s = 8; s >= 1; s >>= 1) x = fmaxf(x, __shfl_xor(x, s, 64));
  return x;
}

__global__ __launch_bounds__(256) void k_shape_aabbs(const ncx_shape* __restrict__ shapes, Pool P,
                                                      const ncx_pose* __restrict__ poses, size_t n, float margin,
                                                      float* b0, float* b1, float* b2, float* b3, float* b4,
                                                      float* b5) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;  // no early exit: every lane takes part in the hull loop
  const int lane = __lane_id();
  ShapeD g = {};
  iso m = {};
  f3 mn = f3{0.f, 0.f, 0.f}, mx = f3{0.f, 0.f, 0.f};
  bool hull = false;
  if (live) {
    g = load_shape(shapes, (uint32_t)i);
    m = load_pose(poses, (uint32_t)i);
    switch (g.type) {
      case NCX_BALL:  // ball_aabb (aabb_ball.rs:8-13)
        mn = m.t + f3{-g.a, -g.a, -g.a};
        mx = m.t + f3{g.a, g.a, g.a};
        break;
      case NCX_CUBOID: {  // center +- |R| h (aabb_cuboid.rs:7-14, isometry_ops.rs:43-45)
        float qi = m.q.i, qj = m.q.j, qk = m.q.k, qw = m.q.w;
        float ww = qw * qw, ii = qi * qi, jj = qj * qj, kk = qk * qk;
        float ij = qi * qj * 2.f, wk = qw * qk * 2.f, wj = qw * qj * 2.f;
        float ik = qi * qk * 2.f, jk = qj * qk * 2.f, wi = qw * qi * 2.f;
        float M[3][3] = {{ww + ii - jj - kk, ij - wk, wj + ik},
                         {wk + ij, ww - ii + jj - kk, jk - wi},
                         {ik - wj, wi + jk, ww - ii - jj + kk}};
        f3 ws;
        ws.x = (fabsf(M[0][0]) * g.a + fabsf(M[0][1]) * g.b) + fabsf(M[0][2]) * g.c;
        ws.y = (fabsf(M[1][0]) * g.a + fabsf(M[1][1]) * g.b) + fabsf(M[1][2]) * g.c;
        ws.z = (fabsf(M[2][0]) * g.a + fabsf(M[2][1]) * g.b) + fabsf(M[2][2]) * g.c;
        mn = m.t - ws;
        mx = m.t + ws;
        break;
      }
      case NCX_CAPSULE: {  // support_map_aabb (aabb_utils.rs:9-31)
        for (int d = 0; d < 3; ++d) {
          f3 e = f3{0.f, 0.f, 0.f};
          set_comp(e, d, 1.f);
          set_comp(mx, d, comp(support<1>(g, m, e, P, 0), d));
          set_comp(e, d, -1.f);
          set_comp(mn, d, comp(support<1>(g, m, e, P, 0), d));
        }
        break;
      }
      default:
        hull = true;
        break;
    }
  }
  // point_cloud_aabb (aabb_utils.rs:59-79): the componentwise min and max of the transformed points
  uint64_t hm = __ballot(hull);
  const int grp = lane >> 4, gl = lane & 15;
  while (hm) {
    // group grp takes the grp-th lowest remaining hull lane (none when fewer remain)
    uint64_t t = hm;
#pragma unroll
    for (int q = 0; q < 3; ++q)
      if (q < grp) t &= t - 1;
    const int h = t ? __ffsll((unsigned long long)t) - 1 : lane;
    // shuffles outside any condition: a lane inactive at a shuffle does not supply its value
    const uint32_t off = (uint32_t)__shfl((int)g.off, h, 64);
    const uint32_t cnt_h = (uint32_t)__shfl((int)g.cnt, h, 64);
    const uint32_t cnt = t ? cnt_h : 0u;
    iso mh;
    mh.t = f3{__shfl(m.t.x, h, 64), __shfl(m.t.y, h, 64), __shfl(m.t.z, h, 64)};
    mh.q.i = __shfl(m.q.i, h, 64);
    mh.q.j = __shfl(m.q.j, h, 64);
    mh.q.k = __shfl(m.q.k, h, 64);
    mh.q.w = __shfl(m.q.w, h, 64);
    float lx = __builtin_inff(), ly = __builtin_inff(), lz = __builtin_inff();
    float ux = -__builtin_inff(), uy = -__builtin_inff(), uz = -__builtin_inff();
    for (uint32_t v0 = (uint32_t)gl; v0 < cnt; v0 += 64) {
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t v = v0 + 16u * u < cnt ? v0 + 16u * u : v0;
        const f3 p = xform_pt(mh, f3{P.x[off + v], P.y[off + v], P.z[off + v]});
        lx = fminf(lx, p.x); ly = fminf(ly, p.y); lz = fminf(lz, p.z);
        ux = fmaxf(ux, p.x); uy = fmaxf(uy, p.y); uz = fmaxf(uz, p.z);
      }
    }
    lx = grp16_min(lx); ly = grp16_min(ly); lz = grp16_min(lz);
    ux = grp16_max(ux); uy = grp16_max(uy); uz = grp16_max(uz);
    // a hull lane among the four taken fetches its box from the first lane of its group
    const bool mine = ((hm >> lane) & 1ull) != 0;
    const int rank = mine ? __popcll(hm & ((1ull << lane) - 1ull)) : 4;
    const int src = rank < 4 ? rank * 16 : lane;
    const float rlx = __shfl(lx, src, 64), rly = __shfl(ly, src, 64), rlz = __shfl(lz, src, 64);
    const float rux = __shfl(ux, src, 64), ruy = __shfl(uy, src, 64), ruz = __shfl(uz, src, 64);
    if (rank < 4) {
      mn = f3{rlx, rly, rlz};
      mx = f3{rux, ruy, ruz};
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) hm &= hm - 1;
  }
  if (!live) return;
  if (margin != 0.f) {  // AABB::loosened (aabb.rs:190-199)
    mn = mn + f3{-margin, -margin, -margin};
    mx = mx + f3{margin, margin, margin};
  }
  b0[i] = mn.x; b1[i] = mn.y; b2[i] = mn.z;
  b3[i] = mx.x; b4[i] = mx.y; b5[i] = mx.z;
}

int shapes_aabbs_on(ncx_shapes* s, ncx_ctx* ctx, const ncx_pose* poses, float margin, float* const out[6]) {
  if (s->n == 0) return NCX_OK;
  hipLaunchKernelGGL(k_shape_aabbs, dim3(ceil_div(s->n, 256)), dim3(256), 0, ctx->stream,
                     s->shapes.as<ncx_shape>(), Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>()}, poses,
                     s->n, margin, out[0], out[1], out[2], out[3], out[4], out[5]);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

extern "C" int ncx_shapes_aabbs_dev(ncx_shapes* s, const ncx_pose* poses, float margin, float* const out[6]) {
  NCX_REQUIRE(s && poses && out, "ncx_shapes_aabbs_dev: null argument");
  NCX_REQUIRE(margin >= 0.f, "ncx_shapes_aabbs_dev: the loosening margin must be positive");
  for (int d = 0; d < 6; ++d) NCX_REQUIRE(out[d], "ncx_shapes_aabbs_dev: null output");
  if (s->n == 0) return NCX_OK;
  hipLaunchKernelGGL(k_shape_aabbs, dim3(ceil_div(s->n, 256)), dim3(256), 0, s->ctx->stream,
                     s->shapes.as<ncx_shape>(), Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>()}, poses,
                     s->n, margin, out[0], out[1], out[2], out[3], out[4], out[5]);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}
