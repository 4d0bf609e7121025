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
#include "ncx_contact.h"

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
  int cap_seg;  // contact-manifold mode: a capsule is its core segment, and each capsule of a pair
                // adds its radius to the pair's prediction
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
// A smaller first-tier arena (24 vertices, 80 faces, 2.7 KiB): nearly twice the arenas per CU
// for pairs whose EPA ends within 20 iterations (most polytope pairs); the rest overflow to T2.
#define T1S_VC 24u
#define T1S_FC 80u
#define T1S_SC 32u
#define T1S_STC 32u
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
    float pred = a.prediction;
    if (MODE == 0 && a.cap_seg) {
      // CapsuleShapeManifoldGenerator / CapsuleCapsuleManifoldGenerator
      // (capsule_shape_manifold_generator.rs:34-37, capsule_capsule_manifold_generator.rs:36-40)
      if (c.g1.type == NCX_CAPSULE) { pred = pred + c.g1.b; c.g1.type = SHAPE_SEGMENT; }
      if (c.g2.type == NCX_CAPSULE) { pred = pred + c.g2.b; c.g2.type = SHAPE_SEGMENT; }
    }
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
    PairOut o = contact_pair<G, LDS ? 1 : 2>(c, pred, has_init, idir, a.P, gl, A, has_init ? a.max_iter : 10000, io);
    int extra_iters = 0;
    if (o.has == 3) {
      // Warm start hit its iteration cap: plain GJK (init_dir = None) judges the pair.
      extra_iters = o.gjk_iters;
      o = contact_pair<G, LDS ? 1 : 2>(c, pred, false, idir, a.P, gl, A, 10000, io);
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
                       Pool{S->hx.as<float>(), S->hy.as<float>(), S->hz.as<float>(), S->tri});
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
  s->hf_shape.release(); s->hf_fn.release(); s->hf_floop.release(); s->hf_loop.release();
  s->hf_edir.release(); s->hf_ev.release(); s->hf_vadj.release(); s->hf_adj.release();
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
      for (int v = 0; v < NCX_T1_VARIANTS; ++v) sample_ms[NCX_T1_VARIANTS * b + v] = s->tune_ms[query][b][v];
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

// First-tier kernel variants, the autotuner's choices.  Contact: group width G = 8, 16, 32 lanes
// per pair with the 44-vertex, 144-face LDS arena (variants 0-2), and G = 8, 16 with the
// 24-vertex, 80-face arena (variants 3-4); every variant holds 8 pairs per block.  The wide
// arena's variants differ in occupancy and lanes per pair only; the small arena's hold more pairs
// per CU and send more pairs to the second tier, which the tuner times with them.  Distance
// queries need no EPA arena and use 128-thread variants without LDS with 8, 16, 32, 4 or 2 lanes.
static const int T1_NVARIANTS[2] = {NCX_T1_VARIANTS, NCX_T1_VARIANTS};
// lanes per pair: contact variants 3, 4 are the small-arena ones; distance variants 3, 4 (no
// arena) take 4 and 2 lanes, for pairs whose supports are closed forms or small hulls
static const int T1_VARIANT_G[2][NCX_T1_VARIANTS] = {{8, 16, 32, 8, 16}, {8, 16, 32, 4, 2}};
static const int T1_VARIANT_BS[2][NCX_T1_VARIANTS] = {{64, 128, 256, 64, 128}, {128, 128, 128, 128, 128}};
// twice the resident blocks (LDS: 4 wide-arena or 7 small-arena blocks per CU; VGPRs: 8 waves).
// A middle arena (32 vertices, 112 faces, 8 lanes) was never the tuner's choice on the scene's
// buckets (5.17 / 2.51 / 3.31 ms against 5.04 / 2.13 / 2.81 for the best) and is left out.
static const unsigned T1_BLOCKS_PER_CU[2][NCX_T1_VARIANTS] = {{8, 8, 4, 14, 8}, {8, 8, 8, 8, 8}};

extern "C" int ncx_shapes_set_tuning(ncx_shapes* s, int query, const int32_t* choice) {
  NCX_REQUIRE(s && choice && (query == 0 || query == 1), "ncx_shapes_set_tuning: bad argument");
  for (int b = 0; b < NCX_NBUCKETS; ++b)
    NCX_REQUIRE(choice[b] >= -1 && choice[b] < T1_NVARIANTS[query], "ncx_shapes_set_tuning: no such variant");
  for (int b = 0; b < NCX_NBUCKETS; ++b) s->tune[query][b] = (int8_t)choice[b];
  return NCX_OK;
}

static int launch_t1(int variant, bool distance, hipStream_t stream, const KArgs& a, unsigned blocks) {
  if (distance) {
    switch (variant) {
      case 0: hipLaunchKernelGGL((k_pairs<8, 128, 0u, 0u, 0u, 0u, 1>), dim3(blocks), dim3(128), 0, stream, a); break;
      case 2: hipLaunchKernelGGL((k_pairs<32, 128, 0u, 0u, 0u, 0u, 1>), dim3(blocks), dim3(128), 0, stream, a); break;
      case 3: hipLaunchKernelGGL((k_pairs<4, 128, 0u, 0u, 0u, 0u, 1>), dim3(blocks), dim3(128), 0, stream, a); break;
      case 4: hipLaunchKernelGGL((k_pairs<2, 128, 0u, 0u, 0u, 0u, 1>), dim3(blocks), dim3(128), 0, stream, a); break;
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
      case 3:
        hipLaunchKernelGGL((k_pairs<8, 64, T1S_VC, T1S_FC, T1S_SC, T1S_STC, 0>), dim3(blocks), dim3(64), 0, stream, a);
        break;
      case 4:
        hipLaunchKernelGGL((k_pairs<16, 128, T1S_VC, T1S_FC, T1S_SC, T1S_STC, 0>), dim3(blocks), dim3(128), 0, stream,
                           a);
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
    // Rounds over the variants, so clock ramps and noise fall on every variant alike; the first
    // round warms up, each variant keeps its least time of the others.
    const int nv = T1_NVARIANTS[mode];
    float ms_v[NCX_T1_VARIANTS] = {};
    for (int rep = 0; rep < TUNE_ROUNDS && !rc; ++rep) {
      for (int v = 0; v < nv; ++v) {
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
    for (int v = 0; v < nv && !rc; ++v) {
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
      blocks = std::max(1u, std::min<unsigned>(blocks, ceil_div(npairs * T1_VARIANT_G[mode][v], T1_VARIANT_BS[mode][v])));
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
        blocks = std::max(1u, std::min<unsigned>(blocks, ceil_div(npairs * T1_VARIANT_G[mode][v], T1_VARIANT_BS[mode][v])));
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
                                 float* w2, float* nrm, float* depth, int32_t* iters, bool may_tune, float* fdir,
                                 bool cap_seg) {
  KArgs a = {};
  a.fdir = fdir;
  a.cap_seg = cap_seg ? 1 : 0;
  a.shapes = S->shapes.as<ncx_shape>();
  a.P = Pool{S->hx.as<float>(), S->hy.as<float>(), S->hz.as<float>(), S->tri};
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
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>(), s->tri};
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
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>(), s->tri};
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
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>(), s->tri};
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
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>(), s->tri};
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

// STAGE: the boxes go through the broad phase's stored-box rule instead of to b0..b5 (the scene
// step's staging, fused): `st` holds the stored boxes (six SoA arrays of n), each is replaced by
// the new box loosened by `stage_margin` unless it contains the new box
// (dbvt_broad_phase.rs:325-347), or by the new box itself on the first step; the stored box is
// mirrored as two float4 in `sb`.
template <bool STAGE>
__global__ __launch_bounds__(256) void k_shape_aabbs(const ncx_shape* __restrict__ shapes, Pool P,
                                                      const ncx_pose* __restrict__ poses, size_t n, float margin,
                                                      float* b0, float* b1, float* b2, float* b3, float* b4,
                                                      float* b5, float* __restrict__ st = nullptr,
                                                      float4* __restrict__ sb = nullptr, float stage_margin = 0.f,
                                                      int first = 0) {
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
  if constexpr (STAGE) {
    float b[6] = {mn.x, mn.y, mn.z, mx.x, mx.y, mx.z};
    if (!first) {
      // AABB::contains (aabb.rs:161-163)
      const bool contained = st[0 * n + i] <= b[0] && st[1 * n + i] <= b[1] && st[2 * n + i] <= b[2] &&
                             st[3 * n + i] >= b[3] && st[4 * n + i] >= b[4] && st[5 * n + i] >= b[5];
      if (contained) return;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        b[d] = b[d] + (-stage_margin);
        b[3 + d] = b[3 + d] + stage_margin;
      }
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) st[d * n + i] = b[d];
    sb[2 * i] = make_float4(b[0], b[1], b[2], b[3]);
    sb[2 * i + 1] = make_float4(b[4], b[5], 0.f, 0.f);
  } else {
    b0[i] = mn.x; b1[i] = mn.y; b2[i] = mn.z;
    b3[i] = mx.x; b4[i] = mx.y; b5[i] = mx.z;
  }
}

// The scene step's shape boxes (loosened by `margin`) through the stored-box rule, into the
// stored boxes `st` and their float4 mirror `sb` (see k_shape_aabbs<true>).
int shapes_aabbs_stage_on(ncx_shapes* s, ncx_ctx* ctx, const ncx_pose* poses, float margin, float* st, float4* sb,
                          float stage_margin, bool first) {
  if (s->n == 0) return NCX_OK;
  hipLaunchKernelGGL(k_shape_aabbs<true>, dim3(ceil_div(s->n, 256)), dim3(256), 0, ctx->stream,
                     s->shapes.as<ncx_shape>(), Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>(), s->tri}, poses,
                     s->n, margin, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, st, sb, stage_margin,
                     first ? 1 : 0);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

extern "C" int ncx_shapes_aabbs_dev(ncx_shapes* s, const ncx_pose* poses, float margin, float* const out[6]) {
  NCX_REQUIRE(s && poses && out, "ncx_shapes_aabbs_dev: null argument");
  NCX_REQUIRE(margin >= 0.f, "ncx_shapes_aabbs_dev: the loosening margin must be positive");
  for (int d = 0; d < 6; ++d) NCX_REQUIRE(out[d], "ncx_shapes_aabbs_dev: null output");
  if (s->n == 0) return NCX_OK;
  hipLaunchKernelGGL(k_shape_aabbs<false>, dim3(ceil_div(s->n, 256)), dim3(256), 0, s->ctx->stream,
                     s->shapes.as<ncx_shape>(), Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>(), s->tri}, poses,
                     s->n, margin, out[0], out[1], out[2], out[3], out[4], out[5]);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}
