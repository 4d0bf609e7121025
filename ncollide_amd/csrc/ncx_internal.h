// ncx_internal.h -- shared host/device definitions of libncx (gfx950 only).
//
// Device arithmetic follows ncollide3d's operation order with no FMA
// contraction (the library is built with -ffp-contract=off):
//   dot(a,b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b) = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   normalize  = v / sqrt(dot(v,v))   (IEEE division and square root)
//   q * v      = t = 2 (q.xyz x v); (t q.w + q.xyz x t) + v
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>
#include "../../include/ncx.h"

#define NCX_FLT_EPS 1.1920929e-07f
#define NCX_T1_VARIANTS 5  // first-tier contact kernel variants (the autotuner's choices)
#define NCX_FLT_MAX 3.40282347e+38f

// ------------------------------------------------------------------ errors
void ncx_set_error(const std::string& msg);
#define NCX_HIP_CHECK(expr)                                                                \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) {                                                                \
      ncx_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                    \
      return NCX_E_HIP;                                                                    \
    }                                                                                      \
  } while (0)
#define NCX_REQUIRE(cond, msg)            \
  do {                                    \
    if (!(cond)) {                        \
      ncx_set_error(msg);                 \
      return NCX_E_INVALID;               \
    }                                     \
  } while (0)

struct ncx_ctx {
  int device;
  hipStream_t stream;
  int num_cus;
};

// Device buffer owned by the library (grows, never shrinks).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int reserve(size_t nbytes) {
    if (nbytes <= bytes) return NCX_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t nb = nbytes < 256 ? 256 : nbytes;
    hipError_t e = hipMalloc(&p, nb);
    if (e != hipSuccess) {
      ncx_set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
      return NCX_E_HIP;
    }
    bytes = nb;
    return NCX_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return (T*)p; }
};

// Layout of ContactTiers::ovf (uint32 words): a header, then the pair indices whose EPA outgrew
// the first-tier arena, those that outgrew the second tier, the third, and the fourth (cap
// entries each).
// Header words [0, NCX_OVF_TIER_WORDS) are the tiers' counts (cleared between tuning samples),
// [NCX_OVF_TIER_WORDS, NCX_OVF_COUNTS) the partition's (which the samples read): each vertex-count
// bucket's range in the partitioned list, and each sort key's count and scatter cursor.  From
// NCX_WORK_BASE on come the work counters of the launches (the tuning samples' first tier, the
// second, third and fourth tiers, one per bucket), each on its own cache line.
constexpr uint32_t NCX_OVERLAP_BINS = 8;  // 4 and 16 bins ordered the scene's pairs no better
constexpr uint32_t NCX_NKEYS = 64 * NCX_OVERLAP_BINS;  // sort keys: vertex-count bucket x 16 shape-type pairs x overlap bins
constexpr uint32_t NCX_OVF_T1_COUNT = 0, NCX_OVF_BAD = 1, NCX_OVF_T2_COUNT = 2, NCX_OVF_T3_COUNT = 3,
                   NCX_OVF_T4_COUNT = 4, NCX_OVF_TIER_WORDS = 8, NCX_BKT_RANGE = 8, NCX_KEY_CNT = 16,
                   NCX_KEY_FILL = NCX_KEY_CNT + NCX_NKEYS, NCX_OVF_COUNTS = NCX_KEY_FILL + NCX_NKEYS;
constexpr uint32_t NCX_WORK_SET = 64, NCX_WORK_BASE = NCX_OVF_COUNTS, NCX_WORK_TUNE = 0, NCX_WORK_T2 = 1, NCX_WORK_T3 = 2,
                   NCX_WORK_T4 = 3, NCX_WORK_BKT = 4, NCX_WORK_SETS = 8,
                   NCX_OVF_HEADER = NCX_WORK_BASE + NCX_WORK_SETS * NCX_WORK_SET;
// Pairs are partitioned by their larger hull vertex count into NCX_NBUCKETS buckets, each run by
// the kernel variant the autotuner picked for it.
constexpr int NCX_NBUCKETS = 4;
static_assert(NCX_OVF_T4_COUNT < NCX_OVF_TIER_WORDS && NCX_OVF_TIER_WORDS <= NCX_BKT_RANGE &&
                  NCX_BKT_RANGE + 2 * NCX_NBUCKETS <= NCX_KEY_CNT && NCX_NKEYS == 16 * NCX_OVERLAP_BINS * NCX_NBUCKETS &&
                  NCX_WORK_BKT + NCX_NBUCKETS <= NCX_WORK_SETS,
              "overflow header layout");
// The work counter of launch k.
inline uint32_t* ncx_work_set(uint32_t* ovf, uint32_t k) { return ovf + NCX_WORK_BASE + k * NCX_WORK_SET; }

struct KArgs;  // contact kernel arguments (ncx_narrow.hip)

// Overflow lists (NCX_OVF_* layout) and the arguments of the last tiered contact launch, one per
// launch site (a shape set's batch calls, each scene), so launches on different streams never
// share them.
struct ContactTiers {
  DevBuf ovf;
  DevBuf blist;      // pair indices partitioned by sort key
  DevBuf bkey;       // each pair's sort key (count pass -> scatter pass)
  size_t bcap = 0;   // entries of blist and bkey
  uint32_t cap = 0;  // entries per overflow list
  KArgs* last = nullptr;
  DevBuf sus[4];     // saved EPA runs handed on: T1 -> T2, T2 -> T3, T3 -> T4, T4 -> device-memory tier
  uint32_t* t1_list() const { return ovf.as<uint32_t>() + NCX_OVF_HEADER; }
  uint32_t* t2_list() const { return ovf.as<uint32_t>() + NCX_OVF_HEADER + cap; }
  uint32_t* t3_list() const { return ovf.as<uint32_t>() + NCX_OVF_HEADER + 2 * (size_t)cap; }
  uint32_t* t4_list() const { return ovf.as<uint32_t>() + NCX_OVF_HEADER + 3 * (size_t)cap; }
};

// Device-resident shape set (ncx_shapes_create).
struct ncx_shapes {
  ncx_ctx* ctx;
  size_t n = 0, npts = 0;
  DevBuf shapes;      // ncx_shape[n]
  DevBuf hx, hy, hz;  // hull pool
  ContactTiers tiers; // for ncx_contact_batch / ncx_distance_batch
  DevBuf qerr;        // error flag of the time-of-impact, ray-cast and composite queries
  const float4* tri = nullptr;  // triangles of the Triangle parts (a mesh's combined set only)
  // polygonal faces, edges and adjacency of the hulls (ConvexHull::try_new), built on the first
  // contact-manifold call: per shape (face base, face count, edge base, 0); per face normal and
  // loop run; loop entries (vertex, edge); per edge direction and end vertices; per pool point
  // the adjacency run; adjacency entries (face, edge).  Ids inside a hull are local.
  bool hf_ready = false;
  DevBuf hf_shape, hf_fn, hf_floop, hf_loop, hf_edir, hf_ev, hf_vadj, hf_adj;
  // autotuner choice per query (0 contact, 1 distance) and bucket: an index into the first-tier
  // kernel variants, or -1 before tuning
  int8_t tune[2][NCX_NBUCKETS] = {{-1, -1, -1, -1}, {-1, -1, -1, -1}};
  float tune_ms[2][NCX_NBUCKETS][NCX_T1_VARIANTS] = {};
};

// ------------------------------------------------------------------ device math
struct f3 {
  float x, y, z;
};
__host__ __device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__host__ __device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__host__ __device__ __forceinline__ float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__host__ __device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ __forceinline__ float norm_sq(f3 a) { return dot(a, a); }
__device__ __forceinline__ float norm(f3 a) { return __builtin_sqrtf(norm_sq(a)); }
__device__ __forceinline__ f3 normalize(f3 a) { return a / norm(a); }
// nalgebra Unit::try_new_and_get: fails when norm <= min_norm.
__device__ __forceinline__ bool try_normalize(f3 v, float min_norm, f3* out, float* n_out) {
  float n = norm(v);
  if (n <= min_norm) return false;
  *out = v / n;
  *n_out = n;
  return true;
}
__device__ __forceinline__ float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
__device__ __forceinline__ void set_comp(f3& v, int i, float s) {
  if (i == 0) v.x = s; else if (i == 1) v.y = s; else v.z = s;
}

struct quat {
  float i, j, k, w;
};
__device__ __forceinline__ f3 qrot(quat q, f3 v) {
  f3 qv = f3{q.i, q.j, q.k};
  f3 t = cross(qv, v) * 2.0f;
  f3 c = cross(qv, t);
  return (t * q.w + c) + v;
}
__device__ __forceinline__ f3 qrot_inv(quat q, f3 v) { return qrot(quat{-q.i, -q.j, -q.k, q.w}, v); }

struct iso {
  f3 t;
  quat q;
};
__device__ __forceinline__ f3 xform_pt(const iso& m, f3 p) { return qrot(m.q, p) + m.t; }
__device__ __forceinline__ f3 xform_vec(const iso& m, f3 v) { return qrot(m.q, v); }
__device__ __forceinline__ f3 inv_xform_pt(const iso& m, f3 p) { return qrot_inv(m.q, p - m.t); }
__device__ __forceinline__ f3 inv_xform_vec(const iso& m, f3 v) { return qrot_inv(m.q, v); }
__device__ __forceinline__ iso load_pose(const ncx_pose* poses, uint32_t i) {
  const float4* p = (const float4*)(poses + i);
  float4 a = p[0], b = p[1];
  iso m;
  m.t = f3{a.x, a.y, a.z};
  m.q = quat{b.x, b.y, b.z, b.w};
  return m;
}
__device__ __forceinline__ iso iso_identity() {
  iso m;
  m.t = f3{0.f, 0.f, 0.f};
  m.q = quat{0.f, 0.f, 0.f, 1.f};
  return m;
}

// ------------------------------------------------------------------ LBVH (shared by broad phase and ray casting)
// Internal node, 64 bytes: the boxes of both children inline so one node visit is
// four 16-byte loads.  child_l/child_r: internal index, or (leaf position | LEAF_BIT).
// last_l/last_r: the largest sorted leaf position under each child.
#define NCX_LEAF_BIT 0x80000000u
// In LBVH::parent entries the same bit means "right child of", bit 30 that the parent's leaf
// range spans refit workgroups (set once per topology), and the low 30 bits are the parent.
#define NCX_RIGHT_BIT 0x80000000u
#define NCX_SPAN_BIT 0x40000000u
#define NCX_PARENT_MASK 0x3FFFFFFFu
struct alignas(16) LNode {
  float lminx, lminy, lminz, lmaxx;
  float lmaxy, lmaxz, rminx, rminy;
  float rminz, rmaxx, rmaxy, rmaxz;
  uint32_t child_l, child_r, last_l, last_r;
};

struct LBVH {
  size_t n = 0;          // number of leaves
  DevBuf nodes;          // LNode[n-1]
  DevBuf parent;         // uint32 parent of internal nodes [n-1] then leaves [n]; side bit 31, span bit 30
  DevBuf order;          // uint32 sorted position -> primitive index
  DevBuf keys;           // Morton codes (sorted)
  DevBuf tmp_keys, tmp_vals, hist, flags, bounds, scan_tmp;
  DevBuf first;          // uint32 first sorted leaf position under each internal node
  DevBuf upper, upper_off;  // refit stage-2 node list (deepest level first) and level offsets
  DevBuf node_depth;        // depth of each internal node (or LBVH_MAX_DEPTH when not in stage 2)
  void release() {
    first.release(); upper.release(); upper_off.release(); node_depth.release();
    nodes.release(); parent.release(); order.release(); keys.release();
    tmp_keys.release(); tmp_vals.release(); hist.release(); flags.release(); bounds.release(); scan_tmp.release();
  }
};

// Depth bound of an LBVH over 30-bit Morton codes with index tie-breaks: along any root-to-leaf
// path the split position (the common prefix length of the 62-bit key:index) strictly grows.
constexpr uint32_t LBVH_MAX_DEPTH = 64;

// Builds the hierarchy over n boxes given as device SoA.  `expand` > 0 grows every
// stored box outward by expand * (1 + |coordinate|) (used for conservative ray culling).
// `aos`, when given, holds the same boxes as two float4 per box (min xyz, max x | max yz, pad),
// which the refit gathers instead of the six SoA arrays.
int lbvh_build(ncx_ctx* ctx, LBVH& bvh, const float* const box[6], size_t n, float expand,
               const float4* aos = nullptr);
// Recomputes every node box for new primitive boxes, keeping order and topology.
int lbvh_refit(ncx_ctx* ctx, LBVH& bvh, const float* const box[6], float expand, const float4* aos = nullptr);

// Fills n uint32 with v on the context stream (a kernel, so it is safe inside graph capture).
int fill_u32(ncx_ctx* ctx, uint32_t* p, size_t n, uint32_t v);
// Per-device address of the allocation-free ray cast's scratch array (ncx_raycast.hip).
int rc_scratch_setup(int device);
// Device-wide exclusive scan of n uint32 (in != out); returns total through *total_dev (device) if non-null.
int scan_exclusive_u32(ncx_ctx* ctx, const uint32_t* in, uint32_t* out, size_t n, DevBuf& tmp, uint32_t* total_dev);
// Stable LSD radix sort of (key, val) uint32 pairs on bits [begin_bit, end_bit).
int radix_sort_u32(ncx_ctx* ctx, uint32_t* keys, uint32_t* vals, size_t n, int begin_bit, int end_bit,
                   DevBuf& tmp_keys, DevBuf& tmp_vals, DevBuf& hist, DevBuf& scan_tmp);
// Stable sort of 64-bit keys (given as hi/lo uint32 arrays) with an optional uint32 payload.
int radix_sort_pairs64(ncx_ctx* ctx, uint32_t* hi, uint32_t* lo, size_t n, int hi_bits, int lo_bits, DevBuf& t1,
                       DevBuf& t2, DevBuf& t3, DevBuf& hist, DevBuf& scan_tmp);

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// TriMesh on the device (ncx_raycast.hip builds it; ncx_composite.hip queries it).
struct ncx_mesh {
  ncx_ctx* ctx;
  size_t npoints = 0, ntri = 0;
  LBVH bvh;
  DevBuf tris;      // float4[3 * ntri] in leaf order: a (w = face id bits), b, c
  DevBuf pts, idx;  // uploaded mesh
  DevBuf overflow;  // spill area for traversal stacks deeper than the LDS part
  DevBuf err;       // error flag (stack exhausted)
  DevBuf work;      // ray counters of the persistent cast (one per part)
  DevBuf qframe;    // float[6]: quantisation origin (3) and step (3)
  DevBuf qnodes4;   // the cast tree: the LBVH with levels merged into 4-wide nodes with 16-bit
                    // quantised child boxes (QNode4), depth-first preorder
  uint32_t levels4 = 1;  // levels of the cast tree (root = level 1)
  bool timing = false;              // record events around the cast kernel (ncx_trimesh_set_timing)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // composite queries against shape sets (ncx_composite.hip): the shapes and the mesh's triangles
  // as one shape set (rebuilt when another shape set comes), poses of both, candidate pairs
  ncx_shapes* comp = nullptr;
  const ncx_shapes* comp_src = nullptr;
  size_t comp_src_n = 0;
  DevBuf cposes, ccnt, coffs, cscan, ctot, ca, cb, chas, cw1, cw2, cn, cdepth, citers;
  void release_composite();
};

// Tiered contact launch over pairs counted on the device (ncx_narrow.hip); no host
// synchronisation unless may_tune (the autotuner's first run, never inside graph capture).
int contact_tiers_reserve(ContactTiers& t, size_t cap);
void contact_tiers_release(ContactTiers& t);
int contact_tiers_reserve_list(ContactTiers& t, size_t npairs);
int contacts_launch_device_count(ncx_shapes* S, ContactTiers& T, ncx_ctx* ctx, const ncx_pose* poses,
                                 const uint32_t* ia, const uint32_t* ib, const uint32_t* n_dev, size_t cap,
                                 float prediction, const float* init_dir, int max_iter, uint8_t* has, float* w1,
                                 float* w2, float* nrm, float* depth, int32_t* iters, bool may_tune,
                                 float* fdir = nullptr, bool cap_seg = false);
// Host side after that launch completed: checks, re-runs what both LDS tiers could not hold,
// and reports the number of pairs that needed more than the first-tier arena.
int contacts_resolve_overflow(ContactTiers& T, ncx_ctx* ctx, size_t* n_large);
