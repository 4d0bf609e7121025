// ncx_raycast.hip -- TriMesh ray casting on the GPU.
//
// The reference (src/query/ray/ray_trimesh.rs:23-50) runs a best-first search
// over a median-split BVT with a BinaryHeap per ray.  Here the triangles are
// put in a binned-SAH binary tree (boxes grown outward by a few ulps so culling is
// conservative), merged into a 4-wide cast tree, and each ray runs a depth-first,
// nearest-child-first closest-hit traversal with a per-thread stack in LDS.  The triangle test is
// ray_intersection_with_triangle (src/query/ray/ray_triangle.rs:32-114) with the
// same operation order, so every per-triangle toi is bit-identical to the
// reference's; the hit kept is the (toi, feature) lexicographic minimum.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <vector>
#include "ncx_internal.h"


// Internal node of the cast tree, 64 bytes (four 16-byte loads per visit): up to four children,
// the binary tree with levels merged.  Child boxes as 16-bit integers on one grid over the root
// box, coordinate = fma(q, step, origin) (one fused operation); each minimum is rounded down and
// each maximum up, checked with that same decode, so every decoded box contains the exact one
// and culling stays conservative (only more boxes pass).  Per coordinate the four children side
// by side: lo[c] holds children 0, 1 of coordinate c in its low and high halves, hi[c] children
// 2, 3; coordinates min x, y, z, max x, y, z.
// child: internal 4-wide node index, a leaf (NCX_LEAF_BIT | first << 2 | count - 1: the
// triangles at sorted positions [first, first + count)), or RC_EMPTY.
#define RC_EMPTY 0xFFFFFFFFu
// Triangles a cast-tree leaf may hold: a subtree of the binary tree with at most this many
// triangles becomes one leaf (its triangles are consecutive in sorted order).
#define RC_LEAF_TRIS 2u
static_assert(RC_LEAF_TRIS >= 1 && RC_LEAF_TRIS <= 4, "two bits of a leaf reference hold its count");
struct alignas(16) QNode4 {
  uint32_t lo[6];
  uint32_t hi[6];
  uint32_t child[4];
};
static_assert(sizeof(QNode4) == 64, "QNode4 is one 64-byte line");

__device__ __forceinline__ float qdec(uint32_t q, float o, float st) { return __builtin_fmaf((float)q, st, o); }

// Quantisation frame: origin = root box minimum, step = extent / 65535 enlarged by 2^-20 so the
// top code decodes at or above the root maximum (the kernel below checks every box).
__global__ void k_qframe(const LNode* __restrict__ nodes, float* __restrict__ frame) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const LNode& r = nodes[0];
  float mn[3] = {fminf(r.lminx, r.rminx), fminf(r.lminy, r.rminy), fminf(r.lminz, r.rminz)};
  float mx[3] = {fmaxf(r.lmaxx, r.rmaxx), fmaxf(r.lmaxy, r.rmaxy), fmaxf(r.lmaxz, r.rmaxz)};
  for (int d = 0; d < 3; ++d) {
    frame[d] = mn[d];
    frame[3 + d] = (mx[d] - mn[d]) / 65535.f * (1.f + 9.5367431640625e-07f);
  }
}

#define RC_BLOCK 128
#define RC_LDS_STACK 24
#define RC_GLOBAL_STACK 104
#define RC_GRID_BLOCKS 4096

__global__ __launch_bounds__(256) void k_tri_reorder(const float* __restrict__ pts, const uint32_t* __restrict__ idx,
                                                      const uint32_t* __restrict__ order, size_t ntri,
                                                      float4* __restrict__ tris) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ntri) return;
  uint32_t f = order[k];
  const float* a = pts + 3 * idx[3 * f];
  const float* b = pts + 3 * idx[3 * f + 1];
  const float* c = pts + 3 * idx[3 * f + 2];
  tris[3 * k + 0] = make_float4(a[0], a[1], a[2], __uint_as_float(f));
  tris[3 * k + 1] = make_float4(b[0], b[1], b[2], 0.f);
  tris[3 * k + 2] = make_float4(c[0], c[1], c[2], 0.f);
}

// ray_intersection_with_triangle (ray_triangle.rs:32-114); false when the ray misses.
__device__ __forceinline__ bool ray_tri(f3 a, f3 b, f3 c, f3 o, f3 dir, float* toi_out, f3* n_out, bool* back) {
  f3 ab = b - a, ac = c - a;
  f3 n = cross(ab, ac);
  float d = dot(n, dir);
  if (d == 0.f) return false;
  f3 ap = o - a;
  float t = dot(ap, n);
  if ((t < 0.f && d < 0.f) || (t > 0.f && d > 0.f)) return false;
  *back = !(d < 0.f);
  d = fabsf(d);
  f3 e = -cross(dir, ap);
  float v, w;
  if (t < 0.f) {
    v = -dot(ac, e);
    if (v < 0.f || v > d) return false;
    w = dot(ab, e);
    if (w < 0.f || v + w > d) return false;
    float invd = 1.f / d;
    *toi_out = -t * invd;
    *n_out = -normalize(n);
  } else {
    v = dot(ac, e);
    if (v < 0.f || v > d) return false;
    w = -dot(ab, e);
    if (w < 0.f || v + w > d) return false;
    float invd = 1.f / d;
    *toi_out = t * invd;
    *n_out = normalize(n);
  }
  return true;
}

// Slab test of one box; axes with a zero direction component use the
// containment test of AABB::toi_with_ray (ray_aabb.rs:20-23).  Returns the entry
// parameter, or NaN on a miss (so every `t <= x` test on it is false).  Written without
// branches: the interval of an axis with a zero component is left unchanged by a select, and
// its containment test only decides the miss, so the result is the same as returning early.
__device__ __forceinline__ float slab(float mnx, float mny, float mnz, float mxx, float mxy, float mxz, f3 o, f3 inv,
                                      bool zx, bool zy, bool zz, float tmax0) {
  const float ax = (mnx - o.x) * inv.x, bx = (mxx - o.x) * inv.x;
  const float ay = (mny - o.y) * inv.y, by = (mxy - o.y) * inv.y;
  const float az = (mnz - o.z) * inv.z, bz = (mxz - o.z) * inv.z;
  const bool miss = (zx & ((o.x < mnx) | (o.x > mxx))) | (zy & ((o.y < mny) | (o.y > mxy))) |
                    (zz & ((o.z < mnz) | (o.z > mxz)));
  float tmin = 0.f, tmax = tmax0;
  tmin = zx ? tmin : fmaxf(tmin, fminf(ax, bx));
  tmax = zx ? tmax : fminf(tmax, fmaxf(ax, bx));
  tmin = zy ? tmin : fmaxf(tmin, fminf(ay, by));
  tmax = zy ? tmax : fminf(tmax, fmaxf(ay, by));
  tmin = zz ? tmin : fmaxf(tmin, fminf(az, bz));
  tmax = zz ? tmax : fminf(tmax, fmaxf(az, bz));
  return (!miss & (tmin <= tmax)) ? tmin : __builtin_nanf("");
}

// slab for a ray with no zero direction component: the same operations in the same order
// without the containment tests.
__device__ __forceinline__ float slab_nz(float mnx, float mny, float mnz, float mxx, float mxy, float mxz, f3 o, f3 inv,
                                         float tmax0) {
  const float ax = (mnx - o.x) * inv.x, bx = (mxx - o.x) * inv.x;
  const float ay = (mny - o.y) * inv.y, by = (mxy - o.y) * inv.y;
  const float az = (mnz - o.z) * inv.z, bz = (mxz - o.z) * inv.z;
  float tmin = fmaxf(0.f, fminf(ax, bx)), tmax = fminf(tmax0, fmaxf(ax, bx));
  tmin = fmaxf(tmin, fminf(ay, by));
  tmax = fminf(tmax, fmaxf(ay, by));
  tmin = fmaxf(tmin, fminf(az, bz));
  tmax = fminf(tmax, fmaxf(az, bz));
  return tmin <= tmax ? tmin : __builtin_nanf("");
}

template <bool STATS>
__global__ __launch_bounds__(RC_BLOCK) void k_cast(const LNode* __restrict__ nodes, const float4* __restrict__ tris,
                                                    uint32_t nleaves, uint32_t ntri_total, ncx_pose pose,
                                                    const float* __restrict__ ox, const float* __restrict__ oy,
                                                    const float* __restrict__ oz, const float* __restrict__ dx,
                                                    const float* __restrict__ dy, const float* __restrict__ dz,
                                                    size_t nrays, float max_toi, float* __restrict__ toi_out,
                                                    float* __restrict__ nrm_out, uint32_t* __restrict__ feat_out,
                                                    uint32_t* __restrict__ stat_nodes, uint32_t* __restrict__ stat_tris,
                                                    uint32_t* __restrict__ ovf, uint32_t* __restrict__ err) {
  __shared__ uint32_t lstack[RC_LDS_STACK * RC_BLOCK];
  const int tid = threadIdx.x;
  const size_t gtid = (size_t)blockIdx.x * RC_BLOCK + tid;
  uint32_t* gstack = ovf + gtid * RC_GLOBAL_STACK;
  iso m;
  m.t = f3{pose.tx, pose.ty, pose.tz};
  m.q = quat{pose.qi, pose.qj, pose.qk, pose.qw};
  for (size_t r = gtid; r < nrays; r += (size_t)gridDim.x * RC_BLOCK) {
    // Ray::inverse_transform_by (ray.rs:36-41)
    f3 o = inv_xform_pt(m, f3{ox[r], oy[r], oz[r]});
    f3 d = inv_xform_vec(m, f3{dx[r], dy[r], dz[r]});
    bool zx = d.x == 0.f, zy = d.y == 0.f, zz = d.z == 0.f;
    f3 inv = f3{1.f / d.x, 1.f / d.y, 1.f / d.z};
    float best_t = max_toi;
    uint32_t best_f = 0xFFFFFFFFu;
    f3 best_n = f3{0.f, 0.f, 0.f};
    uint32_t n_nodes = 0, n_tris = 0;
    auto test_leaf = [&](uint32_t k) {
      float4 A = tris[3 * k], B = tris[3 * k + 1], C = tris[3 * k + 2];
      float t;
      f3 nn;
      bool back;
      if (STATS) n_tris++;
      if (ray_tri(f3{A.x, A.y, A.z}, f3{B.x, B.y, B.z}, f3{C.x, C.y, C.z}, o, d, &t, &nn, &back) && t <= max_toi) {
        uint32_t f = __float_as_uint(A.w);
        uint32_t feat = back ? f + ntri_total : f;
        if (t < best_t || (t == best_t && feat < best_f)) {
          best_t = t;
          best_f = feat;
          best_n = nn;
        }
      }
    };
    if (nleaves == 1) {
      test_leaf(0);
    } else {
      int sp = 0;
      uint32_t node = 0;
      bool running = true;
      while (running) {
        const float4* q = (const float4*)(nodes + node);
        float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
        if (STATS) n_nodes++;
        uint32_t cl = __float_as_uint(q3.x), cr = __float_as_uint(q3.y);
        float tl = slab(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, o, inv, zx, zy, zz, max_toi);
        float tr = slab(q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, o, inv, zx, zy, zz, max_toi);
        bool hl = tl <= best_t, hr = tr <= best_t;
        if (hl && (cl & NCX_LEAF_BIT)) { test_leaf(cl & ~NCX_LEAF_BIT); hl = false; }
        if (hr && (cr & NCX_LEAF_BIT)) { test_leaf(cr & ~NCX_LEAF_BIT); hr = false; }
        if (hl && hr) {
          uint32_t nearc = tl <= tr ? cl : cr, farc = tl <= tr ? cr : cl;
          if (sp < RC_LDS_STACK) lstack[sp * RC_BLOCK + tid] = farc;
          else if (sp < RC_LDS_STACK + RC_GLOBAL_STACK) gstack[sp - RC_LDS_STACK] = farc;
          else { atomicOr(err, 1u); running = false; }
          sp++;
          node = nearc;
        } else if (hl) {
          node = cl;
        } else if (hr) {
          node = cr;
        } else {
          if (sp == 0) { running = false; }
          else {
            sp--;
            node = sp < RC_LDS_STACK ? lstack[sp * RC_BLOCK + tid] : gstack[sp - RC_LDS_STACK];
          }
        }
      }
    }
    if (STATS) {
      stat_nodes[r] = n_nodes;
      stat_tris[r] = n_tris;
    } else {
      if (best_f != 0xFFFFFFFFu) {
        f3 wn = xform_vec(m, best_n);  // res.normal = m * res.normal (ray_trimesh.rs:47)
        toi_out[r] = best_t;
        nrm_out[3 * r] = wn.x;
        nrm_out[3 * r + 1] = wn.y;
        nrm_out[3 * r + 2] = wn.z;
      } else {
        toi_out[r] = -1.f;
        nrm_out[3 * r] = 0.f;
        nrm_out[3 * r + 1] = 0.f;
        nrm_out[3 * r + 2] = 0.f;
      }
      feat_out[r] = best_f;
    }
  }
}

// Persistent closest-hit cast.  The per-ray traversal is k_cast's, one node per loop
// iteration; what changes is the scheduling.  k_cast keeps a wavefront busy until its slowest
// ray ends (a 64-ray wave of this workload runs ~3x the mean node count).  Here a lane that
// finishes its ray takes the next one at once from its wavefront's chunk of RC_CHUNK ray
// indices: each iteration the lanes without a ray are counted by a ballot and take consecutive
// indices by their rank among them (mbcnt).  A wavefront whose chunk runs out takes the next
// chunk from the work counters with one atomicAdd, so a counter sees one atomic per RC_CHUNK
// rays (one per ray serialises on that address and runs 8x slower).  The wave exits when no
// lane holds a ray and every counter is past its part of the batch.  Leaf tests are postponed
// to leaf phases shared by the wavefront (below).
#define RC_CHUNK 64u
#define RC_PARTS 8u
#define RC_PART_STRIDE 64u
#define NONE_LEAF 0xFFFFFFFFu
#define RC_LEAF_PHASE 32u
static_assert(RC_LEAF_PHASE >= 1 && RC_LEAF_PHASE <= 64, "leaf phase threshold out of range");
// Finished rays are staged in LDS by chunk and written out a chunk at a time.
#define RC_WAVES (RC_BLOCK / 64)
#define RC_CAST_STACK 12
#define RC_SLOTS 3u
#define RC_STG_BASE (5u * RC_CHUNK + RC_CHUNK / 4u)
#define RC_STG_WORDS (RC_STG_BASE + 1u)
static_assert(RC_CHUNK == 64u, "one staged entry per lane of a chunk");
static_assert(RC_SLOTS >= 2u && RC_SLOTS <= 4u, "slot state is 16 bits a slot in one 64-bit word");
__global__ __launch_bounds__(RC_BLOCK) void k_cast_wide(const QNode4* __restrict__ nodes, const float* __restrict__ qframe,
                                                       const float4* __restrict__ tris,
                                                       uint32_t nleaves, uint32_t ntri_total, ncx_pose pose,
                                                       const float* __restrict__ ox, const float* __restrict__ oy,
                                                       const float* __restrict__ oz, const float* __restrict__ dx,
                                                       const float* __restrict__ dy, const float* __restrict__ dz,
                                                       uint32_t nrays, float max_toi, float* __restrict__ toi_out,
                                                       float* __restrict__ nrm_out, uint32_t* __restrict__ feat_out,
                                                       uint32_t* __restrict__ ovf, uint32_t gdepth,
                                                       uint32_t* __restrict__ err, uint32_t* __restrict__ work) {
  __shared__ uint32_t lstack[RC_CAST_STACK * RC_BLOCK];
  // Result staging: RC_SLOTS chunk slots per wave.  A lane that finishes a ray of a staged chunk
  // puts (toi, normal, feature) at the ray's place in its slot; once every ray of the chunk is
  // in, the wave writes the slot out with whole-line stores (a scattered 4-byte store per
  // finished ray left most output lines partly written in L2, and their write-backs came to
  // ~5x the output bytes).  Chunk k (k = 0, 1, ... in the order the wave takes them) goes into
  // slot k % RC_SLOTS; what chunk k - RC_SLOTS left there unfinished is written out as it
  // stands, and that chunk's rays still running write their results directly.
  // A slot is RC_STG_WORDS words: toi [0, 64), normals [64, 256), features [256, 320), one byte
  // per entry set once it is staged, then the chunk's first ray.
  __shared__ uint32_t stg[RC_WAVES * RC_SLOTS * RC_STG_WORDS];
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const uint32_t wv = (uint32_t)tid / 64u;
  const size_t gtid = (size_t)blockIdx.x * RC_BLOCK + tid;
  uint32_t* gstack = ovf + gtid * gdepth;
  iso m;
  m.t = f3{pose.tx, pose.ty, pose.tz};
  m.q = quat{pose.qi, pose.qj, pose.qk, pose.qw};
  uint32_t r = 0;
  bool has_ray = false, exhausted = false;
  f3 o = f3{0.f, 0.f, 0.f}, d = o, inv = o, best_n = o;
  bool zx = false, zy = false, zz = false;
  float best_t = max_toi;
  uint32_t best_f = 0xFFFFFFFFu, node = 0, top = 0;
  int sp = 0;
  bool top_valid = false;
  uint32_t pl0 = NONE_LEAF, pl1 = NONE_LEAF, pl2 = NONE_LEAF, pl3 = NONE_LEAF;  // leaves hit, not yet tested
  bool fin = false;                          // the ray's traversal has ended
  // The quantisation frame exists only for meshes with internal nodes (more than one triangle).
  f3 qo = f3{0.f, 0.f, 0.f}, qs = qo;
  if (nleaves > 1) {
    qo = f3{qframe[0], qframe[1], qframe[2]};
    qs = f3{qframe[3], qframe[4], qframe[5]};
  }
  // Work distribution: the rays are split into RC_PARTS contiguous parts, each with its own
  // counter on its own cache line (RC_PART_STRIDE words apart).  Atomics on one line serialise:
  // with one counter shared by every wave the cast ran 1.35x longer, and with the RC_PARTS
  // counters packed into one line 2.6x longer.  A wave starts on part w % RC_PARTS and moves on
  // to the next parts as each is used up.
  const uint32_t wave = blockIdx.x * (RC_BLOCK / 64) + (uint32_t)tid / 64;
  uint32_t next = 0, end = 0;  // the wavefront's chunk [next, end) (the same in every lane)
  bool pool_empty = false;
  const uint32_t part0 = wave % RC_PARTS;
  uint32_t part_k = 0;
  // Slot state (the same in every lane), 16 bits a slot: the chunk's length (bits 0-7, 0 = free)
  // and the count of its rays not yet staged (bits 8-15).  Chunk k is still staged while at most
  // k + RC_SLOTS chunks have been taken.
  uint64_t st = 0;
  uint32_t chunks = 0, cb = 0;  // chunks taken, first ray of the current chunk
  uint32_t myq = 0, me = 0;     // the lane's ray: its chunk number, its entry | slot << 6
  bool staged = false;          // the lane staged a result since the last count
  // Writes slot s out: every entry, or (partial) only those already staged.
  auto flush = [&](uint32_t s, bool partial) {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint32_t* sl = stg + (wv * RC_SLOTS + s) * RC_STG_WORDS;
    const uint32_t base = sl[RC_STG_BASE], len = (uint32_t)(st >> (16u * s)) & 0xFFu;
    const uint8_t* ok = (const uint8_t*)(sl + 5 * RC_CHUNK);
    if ((uint32_t)lane < len && (!partial || ok[lane])) {
      toi_out[base + lane] = __uint_as_float(sl[lane]);
      feat_out[base + lane] = sl[4 * RC_CHUNK + lane];
    }
#pragma unroll
    for (uint32_t c = 0; c < 3; ++c) {
      const uint32_t x = c * RC_CHUNK + (uint32_t)lane;
      if (x < 3u * len && (!partial || ok[x / 3u])) nrm_out[3u * (size_t)base + x] = __uint_as_float(sl[RC_CHUNK + x]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  };
  for (;;) {
    // a ray whose traversal and leaf tests are over is finished here, in one place
    auto finish = [&]() {
      float t = -1.f;
      f3 wn = f3{0.f, 0.f, 0.f};
      if (best_f != 0xFFFFFFFFu) {
        wn = xform_vec(m, best_n);  // res.normal = m * res.normal (ray_trimesh.rs:47)
        t = best_t;
      }
      if (chunks <= myq + RC_SLOTS) {
        uint32_t* sl = stg + (wv * RC_SLOTS + (me >> 6)) * RC_STG_WORDS;
        const uint32_t e = me & 63u;
        sl[e] = __float_as_uint(t);
        sl[RC_CHUNK + 3u * e] = __float_as_uint(wn.x);
        sl[RC_CHUNK + 3u * e + 1] = __float_as_uint(wn.y);
        sl[RC_CHUNK + 3u * e + 2] = __float_as_uint(wn.z);
        sl[4 * RC_CHUNK + e] = best_f;
        ((uint8_t*)(sl + 5 * RC_CHUNK))[e] = 1;
        staged = true;
      } else {
        toi_out[r] = t;
        nrm_out[3 * (size_t)r] = wn.x;
        nrm_out[3 * (size_t)r + 1] = wn.y;
        nrm_out[3 * (size_t)r + 2] = wn.z;
        feat_out[r] = best_f;
      }
      has_ray = false;
    };
    if (has_ray && fin && pl0 == NONE_LEAF) finish();
    // count the results staged; write out the chunks now complete
    {
      if (__ballot(staged)) {
#pragma unroll
        for (uint32_t s = 0; s < RC_SLOTS; ++s)
          st -= (uint64_t)__popcll(__ballot(staged && (me >> 6) == s)) << (16u * s + 8u);
        staged = false;
#pragma unroll
        for (uint32_t s = 0; s < RC_SLOTS; ++s) {
          const uint32_t v = (uint32_t)(st >> (16u * s)) & 0xFFFFu;
          if (v != 0 && (v >> 8) == 0) {
            flush(s, false);
            st &= ~(0xFFFFull << (16u * s));
          }
        }
      }
    }
    // refill lanes without a ray
    const bool want = !has_ray && !exhausted;
    const uint64_t need = __ballot(want);
    if (need) {
      if (next >= end && !pool_empty) {
        const int leader = __ffsll((unsigned long long)need) - 1;
        while (part_k < RC_PARTS) {
          const uint32_t p = (part0 + part_k) % RC_PARTS;
          const uint32_t lo = (uint32_t)(((uint64_t)nrays * p) / RC_PARTS);
          const uint32_t hi = (uint32_t)(((uint64_t)nrays * (p + 1)) / RC_PARTS);
          uint32_t base = 0;
          if (lane == leader) base = atomicAdd(work + p * RC_PART_STRIDE, RC_CHUNK);
          base = __builtin_amdgcn_readlane(base, leader);  // wave-uniform: chunk and slot state stay scalar
          if (base < hi - lo) {
            next = lo + base;
            end = hi - next > RC_CHUNK ? next + RC_CHUNK : hi;
            // the chunk goes into the other slot; what is still there is written out as it
            // stands, and that chunk's remaining rays write their results directly
            const uint32_t ns = chunks % RC_SLOTS;
            if ((st >> (16u * ns)) & 0xFFFFu) flush(ns, true);
            const uint64_t len = end - next;
            st = (st & ~(0xFFFFull << (16u * ns))) | ((len | len << 8) << (16u * ns));
            uint32_t* sl = stg + (wv * RC_SLOTS + ns) * RC_STG_WORDS;
            ((uint8_t*)(sl + 5 * RC_CHUNK))[lane] = 0;
            if (lane == 0) sl[RC_STG_BASE] = next;
            cb = next;
            ++chunks;
            break;
          }
          ++part_k;
        }
        if (part_k >= RC_PARTS) pool_empty = true;
      }
      const uint32_t avail = end - next;
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
      if (want) {
        if (rank < avail) {
          r = next + rank;
          myq = chunks - 1u;
          me = (r - cb) | (myq % RC_SLOTS) << 6;
          // Ray::inverse_transform_by (ray.rs:36-41)
          o = inv_xform_pt(m, f3{ox[r], oy[r], oz[r]});
          d = inv_xform_vec(m, f3{dx[r], dy[r], dz[r]});
          zx = d.x == 0.f; zy = d.y == 0.f; zz = d.z == 0.f;
          inv = f3{1.f / d.x, 1.f / d.y, 1.f / d.z};
          best_t = max_toi;
          best_f = 0xFFFFFFFFu;
          best_n = f3{0.f, 0.f, 0.f};
          node = 0;
          sp = 0;
          top_valid = false;
          pl0 = NONE_LEAF;
          pl1 = NONE_LEAF;
          pl2 = NONE_LEAF;
          pl3 = NONE_LEAF;
          fin = false;
          has_ray = true;
        } else if (pool_empty) {
          exhausted = true;
        }
      }
      const uint32_t cnt = (uint32_t)__popcll(need);
      next += cnt < avail ? cnt : avail;
    }
    auto test_leaf = [&](uint32_t ref) {  // the leaf's triangles [ref >> 2, (ref >> 2) + (ref & 3) + 1)
      const uint32_t k0 = ref >> 2, nk = (ref & 3u) + 1u;
#pragma unroll
      for (uint32_t j = 0; j < RC_LEAF_TRIS; ++j) {
        if (j >= nk) break;
        const uint32_t k = k0 + j;
        float4 A = tris[3 * k], B = tris[3 * k + 1], C = tris[3 * k + 2];
        float t;
        f3 nn;
        bool back;
        if (ray_tri(f3{A.x, A.y, A.z}, f3{B.x, B.y, B.z}, f3{C.x, C.y, C.z}, o, d, &t, &nn, &back) && t <= max_toi) {
          uint32_t f = __float_as_uint(A.w);
          uint32_t feat = back ? f + ntri_total : f;
          if (t < best_t || (t == best_t && feat < best_f)) {
            best_t = t;
            best_f = feat;
            best_n = nn;
          }
        }
      }
    };
    // Leaf phase: lanes that hold untested leaves (or whose traversal has ended) wait; once at
    // least RC_LEAF_PHASE / 64 of the lanes with a ray wait, the wavefront tests every held leaf
    // at once, so the triangle loads and tests run for many lanes together rather than in most
    // iterations for a few.
    const bool waiting = has_ray && (pl0 != NONE_LEAF || fin);
    const uint64_t act = __ballot(has_ray);
    if (act == 0) break;
    const uint64_t wt = __ballot(waiting);
    const bool leaf_phase = (uint32_t)__popcll(wt) * 64u >= (uint32_t)__popcll(act) * RC_LEAF_PHASE;
    if (!has_ray) continue;
    if (leaf_phase) {
      if (waiting) {
        if (pl0 != NONE_LEAF) test_leaf(pl0);
        if (pl1 != NONE_LEAF) test_leaf(pl1);
        if (pl2 != NONE_LEAF) test_leaf(pl2);
        if (pl3 != NONE_LEAF) test_leaf(pl3);
        pl0 = NONE_LEAF;
        pl1 = NONE_LEAF;
        pl2 = NONE_LEAF;
        pl3 = NONE_LEAF;
      }
      continue;
    }
    if (waiting) continue;
    if (nleaves == 1) {
      pl0 = 0;  // the one triangle: first 0, count 1
      fin = true;
      continue;
    }
    // One traversal step over a 4-wide node.  The order in which leaves are tested does not
    // change the result: the kept hit is the least (toi, feature) over every tested triangle, and
    // a box is skipped only when its entry parameter exceeds the current best toi.
    const uint4* q = (const uint4*)(nodes + node);
    const uint4 qa = q[0], qb = q[1], qc = q[2], qd = q[3];
    // lo[0..5] = qa.x qa.y qa.z qa.w qb.x qb.y, hi[0..5] = qb.z qb.w qc.x qc.y qc.z qc.w
    const uint32_t lo[6] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y};
    const uint32_t hi[6] = {qb.z, qb.w, qc.x, qc.y, qc.z, qc.w};
    const uint32_t ch[4] = {qd.x, qd.y, qd.z, qd.w};
    const bool nz = __ballot(zx | zy | zz) == 0;  // no lane's ray has a zero direction component
    float tk[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t* w = c < 2 ? lo : hi;
      const int sh = (c & 1) * 16;
      const float x0 = qdec((w[0] >> sh) & 0xFFFFu, qo.x, qs.x), y0 = qdec((w[1] >> sh) & 0xFFFFu, qo.y, qs.y);
      const float z0 = qdec((w[2] >> sh) & 0xFFFFu, qo.z, qs.z), x1 = qdec((w[3] >> sh) & 0xFFFFu, qo.x, qs.x);
      const float y1 = qdec((w[4] >> sh) & 0xFFFFu, qo.y, qs.y), z1 = qdec((w[5] >> sh) & 0xFFFFu, qo.z, qs.z);
      tk[c] = nz ? slab_nz(x0, y0, z0, x1, y1, z1, o, inv, max_toi)
                 : slab(x0, y0, z0, x1, y1, z1, o, inv, zx, zy, zz, max_toi);
    }
    // leaves hit wait for the leaf phase; internal children hit, nearest first
    const float INF = __builtin_inff();
    float k0 = INF, k1 = INF, k2 = INF, k3 = INF;
    uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t cr = ch[c];
      const bool hit = cr != RC_EMPTY && tk[c] <= best_t;
      if (hit && (cr & NCX_LEAF_BIT)) {
        const uint32_t lf = cr & ~NCX_LEAF_BIT;
        if (pl0 == NONE_LEAF) pl0 = lf;
        else if (pl1 == NONE_LEAF) pl1 = lf;
        else if (pl2 == NONE_LEAF) pl2 = lf;
        else pl3 = lf;
      } else if (hit) {
        if (c == 0) { k0 = tk[c]; r0 = cr; }
        else if (c == 1) { k1 = tk[c]; r1 = cr; }
        else if (c == 2) { k2 = tk[c]; r2 = cr; }
        else { k3 = tk[c]; r3 = cr; }
      }
    }
    // sorting network on (entry parameter, child): (0,1) (2,3) (0,2) (1,3) (1,2)
    auto cswap = [](float& ka, uint32_t& ra, float& kb, uint32_t& rb) {
      const bool sw = kb < ka;
      const float kt = sw ? kb : ka;
      kb = sw ? ka : kb;
      ka = kt;
      const uint32_t rt = sw ? rb : ra;
      rb = sw ? ra : rb;
      ra = rt;
    };
    cswap(k0, r0, k1, r1);
    cswap(k2, r2, k3, r3);
    cswap(k0, r0, k2, r2);
    cswap(k1, r1, k3, r3);
    cswap(k1, r1, k2, r2);
    // The stack's top entry stays in a register; pushing over it moves it to memory.
    auto push = [&](uint32_t x) {
      if (top_valid) {
        if (sp < RC_CAST_STACK) lstack[sp * RC_BLOCK + tid] = top;
        else if (sp < RC_CAST_STACK + (int)gdepth) gstack[sp - RC_CAST_STACK] = top;
        else { atomicOr(err, 1u); fin = true; }
        sp++;
      }
      top = x;
      top_valid = true;
    };
    if (k0 != INF) {
      if (k3 != INF) push(r3);
      if (k2 != INF) push(r2);
      if (k1 != INF) push(r1);
      node = r0;
    } else if (top_valid) {
      node = top;
      top_valid = false;
    } else if (sp == 0) {
      fin = true;
    } else {
      sp--;
      node = sp < RC_CAST_STACK ? lstack[sp * RC_BLOCK + tid] : gstack[sp - RC_CAST_STACK];
    }
  }
}

// Blocks of k_cast_wide resident on one CU (registers and LDS), for the persistent grid.
static unsigned cast_blocks_per_cu() {
  static const unsigned n = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_cast_wide, RC_BLOCK, 0) != hipSuccess || b < 1) b = 1;
    return (unsigned)b;
  }();
  return n;
}

// Quantisation on the host: qdec as the kernel decodes (fmaf is correctly rounded on both sides),
// and the lowest q whose decode is <= x (a minimum) or the highest-needed q whose decode is >= x
// (a maximum).
static inline float qdec_h(uint32_t q, float o, float st) { return std::fmaf((float)q, st, o); }
static inline uint16_t qenc_h(float x, float o, float st, float inv, bool is_max) {
  if (!(st > 0.f)) return 0;
  float t = (x - o) * inv;
  int q = is_max ? (int)std::ceil(t) : (int)std::floor(t);
  q = q < 0 ? 0 : (q > 65535 ? 65535 : q);
  if (is_max) {
    while (q < 65535 && qdec_h((uint32_t)q, o, st) < x) ++q;
  } else {
    while (q > 0 && qdec_h((uint32_t)q, o, st) > x) --q;
  }
  return (uint16_t)q;
}

// The cast tree: each node takes a binary node's two children and, while it has fewer than four,
// replaces the internal child of largest box surface by that child's two children.  Nodes in
// depth-first preorder (a node's first internal child follows it, often in the cache line just
// loaded).  One-time host pass at mesh creation over the exact binary boxes.
static int build_wide(ncx_mesh* m) {
  const size_t ni = m->ntri - 1;
  NCX_REQUIRE(m->ntri < (1ull << 29), "ncx_trimesh_create: more than 2^29 triangles for the cast tree's leaves");
  std::vector<LNode> L(ni);
  std::vector<uint32_t> first(ni);
  float fr[6];
  NCX_HIP_CHECK(hipMemcpy(L.data(), m->bvh.nodes.p, ni * sizeof(LNode), hipMemcpyDeviceToHost));
  NCX_HIP_CHECK(hipMemcpy(first.data(), m->bvh.first.p, ni * 4, hipMemcpyDeviceToHost));
  // a binary child's triangle count (a leaf holds one)
  auto count_of = [&](uint32_t ref) -> uint32_t {
    return (ref & NCX_LEAF_BIT) ? 1u : L[ref].last_r - first[ref] + 1u;
  };
  NCX_HIP_CHECK(hipMemcpy(fr, m->qframe.p, sizeof(fr), hipMemcpyDeviceToHost));
  float inv[3];
  for (int d = 0; d < 3; ++d) inv[d] = fr[3 + d] > 0.f ? 1.f / fr[3 + d] : 0.f;
  struct Ch {
    float b[6];
    uint32_t ref;
  };
  auto area = [](const float* b) {
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return dx * dy + dy * dz + dz * dx;
  };
  std::vector<QNode4> out;
  out.reserve(ni / 2 + 4);
  uint32_t max_depth = 0;
  bool ok = true;
  // explicit stack of (binary node, out index, depth); a node's slot is reserved when it is
  // reached, its children's slots as they are popped (preorder: first child first)
  struct Item {
    uint32_t v, slot_of_parent, cidx, depth;
  };
  std::vector<Item> st;
  st.push_back({0u, 0xFFFFFFFFu, 0u, 0u});
  while (!st.empty()) {
    const Item it = st.back();
    st.pop_back();
    if (it.v >= ni) {
      ncx_set_error("ncx_trimesh_create: malformed tree");
      return NCX_E_INVALID;
    }
    const uint32_t idx = (uint32_t)out.size();
    out.push_back(QNode4{});
    if (it.slot_of_parent != 0xFFFFFFFFu) out[it.slot_of_parent].child[it.cidx] = idx;
    max_depth = std::max(max_depth, it.depth);
    const LNode& a = L[it.v];
    Ch c[4];
    c[0] = Ch{{a.lminx, a.lminy, a.lminz, a.lmaxx, a.lmaxy, a.lmaxz}, a.child_l};
    c[1] = Ch{{a.rminx, a.rminy, a.rminz, a.rmaxx, a.rmaxy, a.rmaxz}, a.child_r};
    int n = 2;
    while (n < 4) {
      int best = -1;
      float ba = -1.f;
      for (int i = 0; i < n; ++i)
        if (count_of(c[i].ref) > RC_LEAF_TRIS && area(c[i].b) > ba) { ba = area(c[i].b); best = i; }
      if (best < 0) break;
      const LNode& u = L[c[best].ref];
      c[best] = Ch{{u.lminx, u.lminy, u.lminz, u.lmaxx, u.lmaxy, u.lmaxz}, u.child_l};
      c[n++] = Ch{{u.rminx, u.rminy, u.rminz, u.rmaxx, u.rmaxy, u.rmaxz}, u.child_r};
    }
    QNode4& q = out[idx];
    for (int k = 0; k < 6; ++k) { q.lo[k] = 0; q.hi[k] = 0; }
    for (int i = 0; i < 4; ++i) {
      if (i >= n) { q.child[i] = RC_EMPTY; continue; }
      for (int k = 0; k < 6; ++k) {
        const int d = k % 3;
        const bool is_max = k >= 3;
        const uint16_t v = qenc_h(c[i].b[k], fr[d], fr[3 + d], inv[d], is_max);
        const float dv = qdec_h(v, fr[d], fr[3 + d]);
        ok = ok && (is_max ? dv >= c[i].b[k] : dv <= c[i].b[k]);
        uint32_t& w = i < 2 ? q.lo[k] : q.hi[k];
        w |= (uint32_t)v << ((i & 1) * 16);
      }
      // leaves: a binary leaf, or a subtree of at most RC_LEAF_TRIS triangles; internal children
      // are patched when they are emitted
      const uint32_t cnt = count_of(c[i].ref);
      if (cnt <= RC_LEAF_TRIS) {
        const uint32_t f0 = (c[i].ref & NCX_LEAF_BIT) ? (c[i].ref & ~NCX_LEAF_BIT) : first[c[i].ref];
        q.child[i] = NCX_LEAF_BIT | f0 << 2 | (cnt - 1u);
      } else {
        q.child[i] = c[i].ref;
      }
    }
    // push internal children in reverse so the first is emitted first
    for (int i = n - 1; i >= 0; --i)
      if (count_of(c[i].ref) > RC_LEAF_TRIS) st.push_back({c[i].ref, idx, (uint32_t)i, it.depth + 1});
  }
  NCX_REQUIRE(ok, "ncx_trimesh_create: a quantised 4-wide box does not contain its exact box");
  m->levels4 = max_depth + 1;
  NCX_REQUIRE(3 * m->levels4 + 1 <= RC_LDS_STACK + RC_GLOBAL_STACK,
              "ncx_trimesh_create: tree too deep for the cast's traversal stack");
  int rc;
  if ((rc = m->qnodes4.reserve(out.size() * sizeof(QNode4)))) return rc;
  NCX_HIP_CHECK(hipMemcpy(m->qnodes4.p, out.data(), out.size() * sizeof(QNode4), hipMemcpyHostToDevice));
  return NCX_OK;
}

// The mesh's binary tree, built on the host by binned SAH (surface area heuristic: a split
// minimises area(left) * count(left) + area(right) * count(right) over 32 centroid bins on each
// axis), one triangle per leaf.  Leaves are numbered in depth-first order, so every internal node
// covers a run [first, last] of sorted positions, as in the device LBVH; node 0 is the root and
// every node's box sits in its parent's child slot, grown by expand * (1 + |coordinate|) at the
// leaves.  Against the Morton-code LBVH it visits fewer nodes per ray on the heightfield workload
// (the cast does not depend on the tree: every result is the least (toi, feature)).  To keep the
// depth within the traversal stacks (the cast tree's 3 entries per level, 128 in all; the
// composite queries' 64 levels), a range whose depth plus log2(size) reaches SAH_MEDIAN_DEPTH is
// split at its median instead, so no leaf is deeper than SAH_MEDIAN_DEPTH + 1.
#define SAH_BINS 32
#define SAH_MEDIAN_DEPTH 40
static void sah_mesh_tree(const float* pts, const uint32_t* idx, size_t nt, float expand, std::vector<LNode>& nodes,
                          std::vector<uint32_t>& order, std::vector<uint32_t>& first) {
  std::vector<float> bl(nt * 3), bh(nt * 3), cen(nt * 3);
  for (size_t t = 0; t < nt; ++t) {
    for (int d = 0; d < 3; ++d) {
      const float a = pts[3 * idx[3 * t] + d], b = pts[3 * idx[3 * t + 1] + d], c = pts[3 * idx[3 * t + 2] + d];
      float lo = std::fmin(std::fmin(a, b), c), hi = std::fmax(std::fmax(a, b), c);
      lo -= expand * (1.f + std::fabs(lo));
      hi += expand * (1.f + std::fabs(hi));
      bl[3 * t + d] = lo;
      bh[3 * t + d] = hi;
      cen[3 * t + d] = 0.5f * lo + 0.5f * hi;
    }
  }
  // a centroid's bin (NaN and out-of-range values go to the end bins, so the index is in range)
  auto bin_of = [](float c, float c0, float scale) {
    const float f = (c - c0) * scale;
    return f >= 0.f ? (f < (float)SAH_BINS ? (int)f : SAH_BINS - 1) : 0;
  };
  // the median split's order: NaN as +infinity, ties by index (a strict weak order)
  auto key = [](float c) { return c == c ? c : INFINITY; };
  std::vector<uint32_t> perm(nt);
  for (size_t t = 0; t < nt; ++t) perm[t] = (uint32_t)t;
  nodes.assign(nt > 1 ? nt - 1 : 0, LNode{});
  order.assign(nt, 0u);
  first.assign(nt > 1 ? nt - 1 : 0, 0u);
  if (nt == 1) return;
  // child references: >= 0 an internal node, < 0 the leaf at sorted position -(ref + 1)
  std::vector<int64_t> cl(nt - 1), cr(nt - 1);
  struct Task {
    uint32_t b, e, depth;
    int64_t node;  // the internal node this range becomes (-1: a leaf)
    int64_t parent;
    int side;
  };
  uint32_t next_node = 0, next_pos = 0;
  std::vector<Task> st;
  st.push_back({0u, (uint32_t)nt, 0u, 0, -1, 0});
  next_node = 1;
  auto area = [](const float* lo, const float* hi) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
  };
  while (!st.empty()) {
    const Task tk = st.back();
    st.pop_back();
    if (tk.e - tk.b == 1) {  // a leaf, numbered in depth-first order
      const uint32_t pos = next_pos++;
      order[pos] = perm[tk.b];
      (tk.side ? cr : cl)[tk.parent] = -(int64_t)pos - 1;
      continue;
    }
    const uint32_t v = (uint32_t)tk.node;
    if (tk.parent >= 0) (tk.side ? cr : cl)[tk.parent] = v;
    const uint32_t n = tk.e - tk.b;
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = tk.b; i < tk.e; ++i)
      for (int d = 0; d < 3; ++d) {
        cmin[d] = std::fmin(cmin[d], cen[3 * perm[i] + d]);
        cmax[d] = std::fmax(cmax[d], cen[3 * perm[i] + d]);
      }
    uint32_t log2n = 0;
    while ((1u << log2n) < n) ++log2n;
    int best_axis = -1, best_bin = 0;
    if (tk.depth + log2n < SAH_MEDIAN_DEPTH) {
      float best = INFINITY;
      for (int d = 0; d < 3; ++d) {
        const float ext = cmax[d] - cmin[d];
        if (!(ext > 0.f)) continue;
        const float scale = SAH_BINS / ext;
        uint32_t cnt[SAH_BINS] = {};
        float lo[SAH_BINS][3], hi[SAH_BINS][3];
        for (int k = 0; k < SAH_BINS; ++k)
          for (int a = 0; a < 3; ++a) { lo[k][a] = INFINITY; hi[k][a] = -INFINITY; }
        for (uint32_t i = tk.b; i < tk.e; ++i) {
          const uint32_t t = perm[i];
          const int k = bin_of(cen[3 * t + d], cmin[d], scale);
          cnt[k]++;
          for (int a = 0; a < 3; ++a) {
            lo[k][a] = std::fmin(lo[k][a], bl[3 * t + a]);
            hi[k][a] = std::fmax(hi[k][a], bh[3 * t + a]);
          }
        }
        float slo[SAH_BINS][3], shi[SAH_BINS][3];
        uint32_t scnt[SAH_BINS];
        for (int k = SAH_BINS - 1; k >= 0; --k)
          for (int a = 0; a < 3; ++a) {
            slo[k][a] = k + 1 < SAH_BINS ? std::fmin(slo[k + 1][a], lo[k][a]) : lo[k][a];
            shi[k][a] = k + 1 < SAH_BINS ? std::fmax(shi[k + 1][a], hi[k][a]) : hi[k][a];
            if (a == 0) scnt[k] = (k + 1 < SAH_BINS ? scnt[k + 1] : 0u) + cnt[k];
          }
        float plo[3] = {INFINITY, INFINITY, INFINITY}, phi[3] = {-INFINITY, -INFINITY, -INFINITY};
        uint32_t pc = 0;
        for (int k = 0; k + 1 < SAH_BINS; ++k) {
          pc += cnt[k];
          for (int a = 0; a < 3; ++a) { plo[a] = std::fmin(plo[a], lo[k][a]); phi[a] = std::fmax(phi[a], hi[k][a]); }
          if (pc == 0 || scnt[k + 1] == 0) continue;
          const float cost = area(plo, phi) * pc + area(slo[k + 1], shi[k + 1]) * scnt[k + 1];
          if (cost < best) { best = cost; best_axis = d; best_bin = k; }
        }
      }
    }
    uint32_t mid;
    if (best_axis >= 0) {
      const float scale = SAH_BINS / (cmax[best_axis] - cmin[best_axis]);
      const float c0 = cmin[best_axis];
      const int ax = best_axis, bb = best_bin;
      mid = (uint32_t)(std::partition(perm.begin() + tk.b, perm.begin() + tk.e,
                                      [&](uint32_t t) { return bin_of(cen[3 * t + ax], c0, scale) <= bb; }) -
                       perm.begin());
    } else {  // no usable split (equal centroids) or depth bound: the median on the longest axis
      int ax = 0;
      for (int d = 1; d < 3; ++d)
        if (cmax[d] - cmin[d] > cmax[ax] - cmin[ax]) ax = d;
      mid = tk.b + n / 2;
      std::nth_element(perm.begin() + tk.b, perm.begin() + mid, perm.begin() + tk.e, [&](uint32_t x, uint32_t y) {
        const float kx = key(cen[3 * x + ax]), ky = key(cen[3 * y + ax]);
        return kx < ky || (kx == ky && x < y);
      });
    }
    const int64_t ln = mid - tk.b > 1 ? (int64_t)next_node++ : -1;
    const int64_t rn = tk.e - mid > 1 ? (int64_t)next_node++ : -1;
    st.push_back({mid, tk.e, tk.depth + 1, rn, (int64_t)v, 1});  // right after left (stack)
    st.push_back({tk.b, mid, tk.depth + 1, ln, (int64_t)v, 0});
  }
  // boxes and leaf ranges, children before parents (children are numbered after their parents)
  std::vector<float> nb(6 * (nt - 1));
  std::vector<uint32_t> nfirst(nt - 1), nlast(nt - 1);
  for (int64_t v = (int64_t)nt - 2; v >= 0; --v) {
    float b[2][6];
    uint32_t f[2], l[2];
    for (int s2 = 0; s2 < 2; ++s2) {
      const int64_t c = s2 ? cr[v] : cl[v];
      if (c < 0) {
        const uint32_t pos = (uint32_t)(-c - 1), t = order[pos];
        for (int a = 0; a < 3; ++a) { b[s2][a] = bl[3 * t + a]; b[s2][3 + a] = bh[3 * t + a]; }
        f[s2] = l[s2] = pos;
      } else {
        for (int a = 0; a < 6; ++a) b[s2][a] = nb[6 * c + a];
        f[s2] = nfirst[c];
        l[s2] = nlast[c];
      }
    }
    for (int a = 0; a < 3; ++a) {
      nb[6 * v + a] = std::fmin(b[0][a], b[1][a]);
      nb[6 * v + 3 + a] = std::fmax(b[0][3 + a], b[1][3 + a]);
    }
    nfirst[v] = f[0];
    nlast[v] = l[1];
    LNode& o = nodes[v];
    o.lminx = b[0][0]; o.lminy = b[0][1]; o.lminz = b[0][2]; o.lmaxx = b[0][3]; o.lmaxy = b[0][4]; o.lmaxz = b[0][5];
    o.rminx = b[1][0]; o.rminy = b[1][1]; o.rminz = b[1][2]; o.rmaxx = b[1][3]; o.rmaxy = b[1][4]; o.rmaxz = b[1][5];
    o.child_l = cl[v] < 0 ? ((uint32_t)(-cl[v] - 1) | NCX_LEAF_BIT) : (uint32_t)cl[v];
    o.child_r = cr[v] < 0 ? ((uint32_t)(-cr[v] - 1) | NCX_LEAF_BIT) : (uint32_t)cr[v];
    o.last_l = l[0];
    o.last_r = l[1];
    first[v] = f[0];
  }
}

extern "C" int ncx_trimesh_create(ncx_ctx* ctx, const float* points_xyz, size_t npoints, const uint32_t* indices,
                                  size_t ntriangles, ncx_mesh** out) {
  NCX_REQUIRE(ctx && out, "ncx_trimesh_create: null argument");
  NCX_REQUIRE(ntriangles > 0 && npoints > 0 && points_xyz && indices, "ncx_trimesh_create: empty mesh");
  NCX_REQUIRE(ntriangles < 0x7FFFFFFFull, "ncx_trimesh_create: too many triangles");
  for (size_t i = 0; i < 3 * ntriangles; ++i)
    NCX_REQUIRE(indices[i] < npoints, "ncx_trimesh_create: triangle index out of range");
  ncx_mesh* m = new ncx_mesh;
  m->ctx = ctx;
  m->npoints = npoints;
  m->ntri = ntriangles;
  int rc = 0;
  auto fail = [&](int code) {
    m->bvh.release();
    m->tris.release(); m->pts.release(); m->idx.release(); m->err.release(); m->work.release();
    m->qframe.release(); m->qnodes4.release();
    delete m;
    return code;
  };
  if ((rc = m->pts.reserve(npoints * 12)) || (rc = m->idx.reserve(ntriangles * 12)) ||
      (rc = m->tris.reserve(ntriangles * 48)) ||
      (rc = m->err.reserve(4)) || (rc = m->work.reserve(4 * RC_PARTS * RC_PART_STRIDE)))
    return fail(rc);
  if (hipMemcpyAsync(m->pts.p, points_xyz, npoints * 12, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipMemcpyAsync(m->idx.p, indices, ntriangles * 12, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
    ncx_set_error("ncx_trimesh_create: upload failed");
    return fail(NCX_E_HIP);
  }
  // Grow node boxes by 2^-18 (1 + |x|): larger than the rounding of any slab or triangle toi.
  {
    std::vector<LNode> nodes;
    std::vector<uint32_t> order, first;
    sah_mesh_tree(points_xyz, indices, ntriangles, 3.8146973e-06f, nodes, order, first);
    m->bvh.n = ntriangles;
    if ((rc = m->bvh.order.reserve(ntriangles * 4)) ||
        (rc = m->bvh.nodes.reserve((ntriangles > 1 ? ntriangles - 1 : 1) * sizeof(LNode))) ||
        (rc = m->bvh.first.reserve((ntriangles > 1 ? ntriangles - 1 : 1) * 4)))
      return fail(rc);
    NCX_HIP_CHECK(hipMemcpy(m->bvh.order.p, order.data(), ntriangles * 4, hipMemcpyHostToDevice));
    if (ntriangles > 1) {
      NCX_HIP_CHECK(hipMemcpy(m->bvh.nodes.p, nodes.data(), nodes.size() * sizeof(LNode), hipMemcpyHostToDevice));
      NCX_HIP_CHECK(hipMemcpy(m->bvh.first.p, first.data(), first.size() * 4, hipMemcpyHostToDevice));
    }
  }
  if ((rc = fill_u32(ctx, m->err.as<uint32_t>(), 1, 0u))) return fail(rc);
  hipLaunchKernelGGL(k_tri_reorder, dim3(ceil_div(ntriangles, 256)), dim3(256), 0, ctx->stream, m->pts.as<float>(),
                     m->idx.as<uint32_t>(), m->bvh.order.as<uint32_t>(), ntriangles, m->tris.as<float4>());
  if ((rc = m->qframe.reserve(64)) || (rc = fill_u32(ctx, m->qframe.as<uint32_t>(), 16, 0u))) return fail(rc);
  if (ntriangles > 1)
    hipLaunchKernelGGL(k_qframe, dim3(1), dim3(64), 0, ctx->stream, m->bvh.nodes.as<LNode>(), m->qframe.as<float>());
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
    ncx_set_error("ncx_trimesh_create: build failed");
    return fail(NCX_E_HIP);
  }
  if (ntriangles > 1 && (rc = build_wide(m))) return fail(rc);
  *out = m;
  return NCX_OK;
}

extern "C" int ncx_trimesh_destroy(ncx_mesh* m) {
  if (!m) return NCX_OK;
  m->bvh.release();
  m->tris.release(); m->pts.release(); m->idx.release(); m->overflow.release(); m->err.release(); m->work.release();
  m->qframe.release(); m->qnodes4.release();
  m->release_composite();
  if (m->ev0) (void)hipEventDestroy(m->ev0);
  if (m->ev1) (void)hipEventDestroy(m->ev1);
  delete m;
  return NCX_OK;
}

static int cast_impl(ncx_mesh* m, const ncx_pose* pose, const float* const r[6], size_t n, float max_toi, float* toi,
                     float* nrm, uint32_t* feat, uint32_t* sn, uint32_t* st, bool stats, bool sync) {
  NCX_REQUIRE(m && pose && r, "ncx_trimesh_cast_rays: null argument");
  NCX_REQUIRE(n <= 0x7FFFFFFFull, "ncx_trimesh_cast_rays: more than 2^31 - 1 rays in one call");
  if (n == 0) return NCX_OK;
  for (int d = 0; d < 6; ++d) NCX_REQUIRE(r[d], "ncx_trimesh_cast_rays: null ray array");
  ncx_ctx* ctx = m->ctx;
  unsigned blocks = ceil_div(n, RC_BLOCK);
  if (blocks > RC_GRID_BLOCKS) blocks = RC_GRID_BLOCKS;
  int rc;
  if ((rc = m->overflow.reserve((size_t)RC_GRID_BLOCKS * RC_BLOCK * RC_GLOBAL_STACK * 4))) return rc;
  if (stats)
    hipLaunchKernelGGL(k_cast<true>, dim3(blocks), dim3(RC_BLOCK), 0, ctx->stream, m->bvh.nodes.as<LNode>(),
                       m->tris.as<float4>(), (uint32_t)m->ntri, (uint32_t)m->ntri, *pose, r[0], r[1], r[2], r[3], r[4],
                       r[5], n, max_toi, toi, nrm, feat, sn, st, m->overflow.as<uint32_t>(), m->err.as<uint32_t>());
  else {
    // persistent grid: as many blocks as are resident.  The spill stacks take the stack
    // entries the LDS does not hold, within the reserved buffer.
    const uint32_t gdepth = RC_LDS_STACK + RC_GLOBAL_STACK - RC_CAST_STACK;
    unsigned pblocks = std::min<unsigned>(RC_GRID_BLOCKS, (unsigned)ctx->num_cus * cast_blocks_per_cu());
    pblocks = std::min<unsigned>(pblocks, (unsigned)(((size_t)RC_GRID_BLOCKS * RC_GLOBAL_STACK) / gdepth));
    NCX_REQUIRE(m->qframe.p && (m->ntri == 1 || m->qnodes4.p), "ncx_trimesh_cast_rays: mesh nodes missing");
    if ((rc = fill_u32(ctx, m->work.as<uint32_t>(), RC_PARTS * RC_PART_STRIDE, 0u))) return rc;
    if (m->timing) NCX_HIP_CHECK(hipEventRecord(m->ev0, ctx->stream));
    hipLaunchKernelGGL(k_cast_wide, dim3(pblocks), dim3(RC_BLOCK), 0, ctx->stream, m->qnodes4.as<QNode4>(),
                       m->qframe.as<float>(), m->tris.as<float4>(), (uint32_t)m->ntri, (uint32_t)m->ntri, *pose, r[0],
                       r[1], r[2], r[3], r[4], r[5], (uint32_t)n, max_toi, toi, nrm, feat, m->overflow.as<uint32_t>(),
                       gdepth, m->err.as<uint32_t>(), m->work.as<uint32_t>());
    NCX_HIP_CHECK(hipGetLastError());
    if (m->timing) NCX_HIP_CHECK(hipEventRecord(m->ev1, ctx->stream));
  }
  NCX_HIP_CHECK(hipGetLastError());
  if (!sync) return NCX_OK;
  return ncx_trimesh_check(m);
}

// Measurement: events around each cast's traversal kernel (k_cast_wide), read back with
// ncx_trimesh_kernel_ms after the cast has completed.
extern "C" int ncx_trimesh_set_timing(ncx_mesh* m, int enable) {
  NCX_REQUIRE(m, "ncx_trimesh_set_timing: null argument");
  if (enable && !m->ev0) {
    NCX_HIP_CHECK(hipEventCreate(&m->ev0));
    NCX_HIP_CHECK(hipEventCreate(&m->ev1));
  }
  m->timing = enable != 0;
  return NCX_OK;
}

extern "C" int ncx_trimesh_kernel_ms(ncx_mesh* m, float* ms) {
  NCX_REQUIRE(m && ms && m->ev0, "ncx_trimesh_kernel_ms: timing not enabled");
  NCX_HIP_CHECK(hipEventSynchronize(m->ev1));
  NCX_HIP_CHECK(hipEventElapsedTime(ms, m->ev0, m->ev1));
  return NCX_OK;
}

// The traversal-stack error flag is sticky: it is cleared at mesh creation and checked here.
extern "C" int ncx_trimesh_check(ncx_mesh* m) {
  NCX_REQUIRE(m, "ncx_trimesh_check: null argument");
  uint32_t err = 0;
  NCX_HIP_CHECK(hipMemcpyAsync(&err, m->err.p, 4, hipMemcpyDeviceToHost, m->ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
  if (err) {
    ncx_set_error("ncx_trimesh_cast_rays: traversal stack exhausted");
    return NCX_E_CAPACITY;
  }
  return NCX_OK;
}

extern "C" int ncx_trimesh_cast_rays(ncx_mesh* m, const ncx_pose* pose, const float* const ray_soa_dev[6], size_t n,
                                     float max_toi, float* toi_dev, float* normal_dev, uint32_t* feature_dev) {
  NCX_REQUIRE(toi_dev && normal_dev && feature_dev, "ncx_trimesh_cast_rays: null output");
  return cast_impl(m, pose, ray_soa_dev, n, max_toi, toi_dev, normal_dev, feature_dev, nullptr, nullptr, false, true);
}

extern "C" int ncx_trimesh_cast_rays_async(ncx_mesh* m, const ncx_pose* pose, const float* const ray_soa_dev[6],
                                           size_t n, float max_toi, float* toi_dev, float* normal_dev,
                                           uint32_t* feature_dev) {
  NCX_REQUIRE(toi_dev && normal_dev && feature_dev, "ncx_trimesh_cast_rays_async: null output");
  return cast_impl(m, pose, ray_soa_dev, n, max_toi, toi_dev, normal_dev, feature_dev, nullptr, nullptr, false, false);
}

extern "C" int ncx_trimesh_cast_stats(ncx_mesh* m, const ncx_pose* pose, const float* const ray_soa_dev[6], size_t n,
                                      float max_toi, uint32_t* nodes_dev, uint32_t* tris_dev) {
  NCX_REQUIRE(nodes_dev && tris_dev, "ncx_trimesh_cast_stats: null output");
  return cast_impl(m, pose, ray_soa_dev, n, max_toi, nullptr, nullptr, nullptr, nodes_dev, tris_dev, true, true);
}

// ------------------------------------------------------------------ allocation-free cast
// Device scratch of ncx_trimesh_cast_rays_noalloc: the work counters of the persistent cast
// (RC_PARTS counters, RC_PART_STRIDE words apart) followed by the traversal spill stacks.  It is
// a file-scope device array, so each GPU the library runs on holds its own copy, at its own
// address, until the process ends.  Host code reaches it only through the per-device address
// that hipGetSymbolAddress returns (the array's name on the host is not a device address).
#define RC_SCRATCH_BYTES (16u << 20)
#define RC_SCRATCH_WORDS (RC_SCRATCH_BYTES / 4u)
#define RC_SCRATCH_MAX_DEVICES 64
static __device__ __align__(256) uint32_t scratch[RC_SCRATCH_WORDS];
static uint32_t* g_scratch_ptr[RC_SCRATCH_MAX_DEVICES];
static std::mutex g_scratch_mu;

// Looks up the scratch address on `device` once; ncx_init calls it with the device current.
int rc_scratch_setup(int device) {
  NCX_REQUIRE(device >= 0 && device < RC_SCRATCH_MAX_DEVICES, "ncx_init: device index beyond 64");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return NCX_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    ncx_set_error(std::string("ncx_init: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    return NCX_E_HIP;
  }
  g_scratch_ptr[device] = (uint32_t*)p;
  return NCX_OK;
}

extern "C" int ncx_trimesh_cast_rays_noalloc(ncx_mesh* m, const ncx_pose* pose, const float* const r[6], size_t n,
                                             float max_toi, float* toi, float* nrm, uint32_t* feat) {
  NCX_REQUIRE(m && pose && r && toi && nrm && feat, "ncx_trimesh_cast_rays_noalloc: null argument");
  NCX_REQUIRE(n <= 0x7FFFFFFFull, "ncx_trimesh_cast_rays_noalloc: more than 2^31 - 1 rays in one call");
  if (n == 0) return NCX_OK;
  for (int d = 0; d < 6; ++d) NCX_REQUIRE(r[d], "ncx_trimesh_cast_rays_noalloc: null ray array");
  ncx_ctx* ctx = m->ctx;
  NCX_REQUIRE(ctx->device >= 0 && ctx->device < RC_SCRATCH_MAX_DEVICES, "ncx_trimesh_cast_rays_noalloc: bad device");
  uint32_t* sp;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    sp = g_scratch_ptr[ctx->device];
  }
  NCX_REQUIRE(sp, "ncx_trimesh_cast_rays_noalloc: scratch not set up on this device (ncx_init)");
  NCX_REQUIRE(m->qframe.p && (m->ntri == 1 || m->qnodes4.p), "ncx_trimesh_cast_rays_noalloc: mesh nodes missing");
  // A traversal keeps at most 3 pending subtrees per level (the children it does not descend
  // into): one in a register, RC_CAST_STACK in LDS, the rest in its spill stack.
  const uint32_t need = 3 * m->levels4 + 1;
  const uint32_t gdepth = need > RC_CAST_STACK ? need - RC_CAST_STACK : 1u;
  const size_t counters = (size_t)RC_PARTS * RC_PART_STRIDE;
  const size_t per_block = (size_t)RC_BLOCK * gdepth;
  unsigned pblocks = std::min<unsigned>(RC_GRID_BLOCKS, (unsigned)ctx->num_cus * cast_blocks_per_cu());
  pblocks = std::min<unsigned>(pblocks, ceil_div(n, RC_BLOCK));
  const size_t fit = (RC_SCRATCH_WORDS - counters) / per_block;
  if (fit < pblocks) pblocks = (unsigned)fit;
  if (pblocks < std::min<unsigned>((unsigned)ctx->num_cus, ceil_div(n, RC_BLOCK))) {
    ncx_set_error("ncx_trimesh_cast_rays_noalloc: the mesh's " + std::to_string(m->levels4) +
                  "-level tree needs " + std::to_string(per_block * 4) + " bytes of spill stack per block; " +
                  std::to_string(RC_SCRATCH_BYTES) + " bytes of scratch do not hold one block per CU");
    return NCX_E_CAPACITY;
  }
  int rc;
  if ((rc = fill_u32(ctx, sp, counters, 0u))) return rc;
  hipLaunchKernelGGL(k_cast_wide, dim3(pblocks), dim3(RC_BLOCK), 0, ctx->stream, m->qnodes4.as<QNode4>(),
                     m->qframe.as<float>(), m->tris.as<float4>(), (uint32_t)m->ntri, (uint32_t)m->ntri, *pose, r[0],
                     r[1], r[2], r[3], r[4], r[5], (uint32_t)n, max_toi, toi, nrm, feat, sp + counters, gdepth,
                     m->err.as<uint32_t>(), sp);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}
