// ncx_world.hip -- one collision step of a whole scene on the GPU, capturable in a hipGraph.
//
// Stands in for one CollisionWorld::update (src/pipeline/world.rs:104-119) restricted to the hot
// path: every shape is a broad-phase proxy; per step
//   1. world AABBs loosened by the query limit (glue/setup.rs:27-28; collision_object.rs:100-113),
//   2. the DBVTBroadPhase stored-box rule: a proxy's box is replaced by box.loosened(margin) only
//      when the new box is not contained in it (dbvt_broad_phase.rs:325-347); the first step
//      stores the boxes as given (create_proxy, :275-280),
//   3. LBVH refit over the stored boxes (rebuilt every `rebuild_every` steps, default 2: the
//      refit tree's query cost grows each step as shapes move, ~2.5x after 7 refits),
//   4. all intersecting stored-box pairs, counted, prefix-summed and written on the device,
//   5. query::contact on every pair (contact_shape_shape.rs:10-48), optionally with the learned
//      init_dir on polytope pairs.
// Steps 1-5 contain no host synchronisation, so a refit step and a rebuild step are each captured
// once in a hipGraph and replayed (the first step, which tunes and allocates, runs eagerly).
#include "ncx_bvh_query.h"

#define WQ_BLOCK 128

struct ncx_mlp;
int mlp_predict_device_count(ncx_mlp* m, ncx_ctx* ctx, const ncx_shape* shapes, const ncx_pose* poses,
                             const float* desc, const uint32_t* ia, const uint32_t* ib, const uint32_t* n_dev,
                             size_t cap, float* dirs);
int shapes_aabbs_stage_on(ncx_shapes* s, ncx_ctx* ctx, const ncx_pose* poses, float margin, float* st, float4* sb,
                          float stage_margin, bool first);
int shape_descriptors_on(ncx_shapes* s, ncx_ctx* ctx, float* desc);

struct ncx_world {
  ncx_ctx* ctx;   // caller context (its stream orders the step with the caller's work)
  ncx_ctx wctx;   // the world's own stream (capturable)
  ncx_shapes* shapes;
  size_t n, cap;
  float margin, prediction;
  int rebuild_every = 2;
  int since_build = 0;
  bool first = true;
  DevBuf stored;
  DevBuf sbox;  // the stored boxes again, two float4 per proxy (min xyz, max x | max yz), for gathers
  LBVH bvh;
  uint32_t spill = 1;
  DevBuf gstack;  // search-stack spill area, `spill` entries per query thread
  DevBuf qwork;   // the query passes' work counters (WQ_PARTS per pass)
  DevBuf lbuf;    // the count pass's first WQ_LOCAL partners per leaf (sorted leaf positions)
  // the count pass's partners past the first WQ_LOCAL: per leaf a directory (ovdir) of chunks of
  // WQ_OCH entries taken from a pool; a leaf that finds no directory, or outgrows one, is
  // listed WQ_NODIR and traversed again by the write pass
  DevBuf ovdir, odirs, ochunks;
  DevBuf relist;  // the leaves marked WQ_NODIR, for the write pass
  uint32_t ndir = 0, nch = 0;
  // counters: [0] pair total, [1] pair overflow, [2] directories, [3] chunks taken, [4] leaves
  // listed for the write pass
  DevBuf cnt, offs, counters;
  DevBuf pa, pb;
  DevBuf has, w1, w2, nrm, depth_out, iters;
  ContactTiers tiers;  // this scene's contact overflow lists
  ncx_mlp* mlp = nullptr;
  DevBuf desc, dirs;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph_rb = nullptr;       // the rebuild step's graph
  hipGraphExec_t exec_rb = nullptr;
  const ncx_pose* rb_poses = nullptr;
  bool rb_valid = false;
  const ncx_pose* graph_poses = nullptr;
  bool graph_valid = false;
  int last_was_graph = 0;
  bool contacts = true;        // run the narrow phase in each step
  bool last_had_contacts = false;
  DevBuf qcnt, qscan, qtot;    // interference queries: per-query counts, scan scratch, total
  DevBuf rkeys, rvals, rtk, rtv, rhist, rscan;  // first-interference rays: sort keys, order, sort scratch
  // the polyhedron manifold generator's last_gjk_dir kept per pair across steps
  bool temporal = false;
  bool synced = true;          // ncx_world_sync has run since the last step
  uint64_t tw_mask = 0;        // hash table slots - 1 (a power of two, at least twice max_pairs)
  DevBuf tw_keys, tw_dirs;     // pair key (UINT64_MAX = empty) and direction (float4) per slot
  DevBuf tw_init, tw_fdir;     // per pair: the looked-up init_dir, the step's result direction
};

// ------------------------------------------------------------------ last_gjk_dir per pair
// ConvexPolyhedronConvexPolyhedronManifoldGenerator keeps, per pair of polyhedra (Cuboid or
// ConvexHull on both sides), the direction of its last GJK/EPA result and starts the next step's
// GJK from it (convex_polyhedron_convex_polyhedron_manifold_generator.rs:24-33, 96-145); the
// generator lives while the broad phase keeps the pair.  Here the directions live in an
// open-addressing table keyed by the pair key: at the start of a step the previous step's pairs
// and result directions are inserted (after the last step's contacts, including those finished
// by the host-driven tier at sync), then each pair of the new step looks its direction up (none
// when the pair is new).
#define TW_EMPTY 0xFFFFFFFFFFFFFFFFull
__device__ __forceinline__ uint64_t tw_hash(uint64_t k) {  // a 64-bit mix (MurmurHash3 finaliser)
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}
__device__ __forceinline__ bool tw_poly(const ncx_shape* __restrict__ sh, uint32_t a, uint32_t b) {
  const uint32_t ta = sh[a].type, tb = sh[b].type;
  return (ta == NCX_CUBOID || ta == NCX_CONVEX_HULL) && (tb == NCX_CUBOID || tb == NCX_CONVEX_HULL);
}

__global__ void k_tw_insert(const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                            const uint32_t* __restrict__ n_dev, size_t cap, const ncx_shape* __restrict__ sh,
                            const float* __restrict__ fdir, unsigned long long* __restrict__ keys,
                            float4* __restrict__ dirs, uint64_t mask) {
  const size_t n = min(cap, (size_t)*n_dev);
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const uint32_t a = pa[k], b = pb[k];
    const float dx = fdir[3 * k], dy = fdir[3 * k + 1], dz = fdir[3 * k + 2];
    if (!tw_poly(sh, a, b) || (dx == 0.f && dy == 0.f && dz == 0.f)) continue;
    const unsigned long long key = ((unsigned long long)a << 32) | b;
    uint64_t h = tw_hash(key) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
      const unsigned long long prev = atomicCAS(&keys[h], TW_EMPTY, key);
      if (prev == TW_EMPTY || prev == key) {
        dirs[h] = make_float4(dx, dy, dz, 0.f);
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

__global__ void k_tw_lookup(const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                            const uint32_t* __restrict__ n_dev, size_t cap, const ncx_shape* __restrict__ sh,
                            const unsigned long long* __restrict__ keys, const float4* __restrict__ dirs,
                            uint64_t mask, float* __restrict__ init) {
  const size_t n = min(cap, (size_t)*n_dev);
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const uint32_t a = pa[k], b = pb[k];
    float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tw_poly(sh, a, b)) {
      const unsigned long long key = ((unsigned long long)a << 32) | b;
      uint64_t h = tw_hash(key) & mask;
      for (uint64_t probe = 0; probe <= mask; ++probe) {
        const unsigned long long kk = keys[h];
        if (kk == key) { d = dirs[h]; break; }
        if (kk == TW_EMPTY) break;
        h = (h + 1) & mask;
      }
    }
    init[3 * k] = d.x;
    init[3 * k + 1] = d.y;
    init[3 * k + 2] = d.z;
  }
}

// The stored boxes are kept twice: SoA (the interface's layout) and as two float4 per proxy,
// which the refit and the pair query gather by sorted leaf: one cache line per box instead of
// six.  The shape-box kernel writes both (k_shape_aabbs<true>).

__device__ __forceinline__ bool wbox_overlap(const float q[6], float mnx, float mny, float mnz, float mxx, float mxy,
                                             float mxz) {
  return mnx <= q[3] && mny <= q[4] && mnz <= q[5] && mxx >= q[0] && mxy >= q[1] && mxz >= q[2];
}

// Self-query of sorted leaf k against leaves at later positions (each pair is found once, by its
// earlier leaf).  Depth-first with the stack's top in a register, the next WQ_LSTACK entries in
// LDS and the rest in a per-thread spill area of `spill` entries (the tree depth bounds it).
//   MODE_COUNT: count the pairs and keep every partner (leaf positions: no dependent gather
//               inside the traversal): the first WQ_LOCAL in lbuf, the rest in chunks of WQ_OCH
//               entries, listed in the leaf's directory of WQ_DIRN chunks.  A leaf that finds the
//               directories used up, or has more than WQ_LOCAL + WQ_DIRN * WQ_OCH partners, is
//               marked WQ_NODIR and listed in relist (the list's length in counters[4]);
//   MODE_WRITE: for the leaves listed in relist, the partners past the first WQ_LOCAL (leaf
//               positions) at their slots offs[k] + index.
// Persistent: a lane that finishes a leaf takes the next one at once from its wavefront's chunk of
// WQ_CHUNK leaves (ballot + rank), and a wavefront whose chunk is used up takes the next chunk of
// its part of the leaves with one atomicAdd on that part's counter (WQ_PARTS counters, each on
// its own cache line, so the atomics do not serialise on one line); a wave that finds its part
// used up moves on to the next parts.  A block starts on part blockIdx % WQ_PARTS: blocks are
// dealt to the 8 XCDs in turn, so each XCD's L2 first serves one contiguous (spatially close)
// eighth of the leaves and the tree nodes above them.
#define WQ_LSTACK 16
#define WQ_LOCAL 512
#define WQ_OCH 256u
#define WQ_DIRN 64u
#define WQ_NODIR 0xFFFFFFFFu
#define WQ_CHUNK 64u
#define WQ_PARTS 8u
#define WQ_PART_STRIDE 64u
#define WQ_BLOCKS_PER_CU 8u
#define MODE_COUNT 0
#define MODE_WRITE 1

template <int MODE>
__global__ __launch_bounds__(WQ_BLOCK) void k_world_query(const LNode* __restrict__ nodes,
                                                           const uint32_t* __restrict__ order,
                                                           const float4* __restrict__ sbox, uint32_t n,
                                                           uint32_t* __restrict__ cnt, const uint32_t* __restrict__ offs,
                                                           size_t cap, uint32_t* __restrict__ pa,
                                                           uint32_t* __restrict__ pb, uint32_t* __restrict__ counters,
                                                           uint32_t* __restrict__ gstack, uint32_t spill,
                                                           uint32_t* __restrict__ lbuf,
                                                           unsigned long long* __restrict__ visits,
                                                           uint32_t* __restrict__ work, uint32_t* __restrict__ ovdir,
                                                           uint32_t* __restrict__ odirs, uint32_t ndir,
                                                           uint32_t* __restrict__ ochunks, uint32_t nch,
                                                           uint32_t* __restrict__ relist) {
  __shared__ uint32_t lstack[WQ_LSTACK * WQ_BLOCK];
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const uint32_t gtid = blockIdx.x * WQ_BLOCK + tid;
  uint32_t* gs = gstack + (size_t)gtid * spill;
  const uint32_t part0 = blockIdx.x % WQ_PARTS;
  uint32_t part_k = 0;
  uint32_t next = 0, end = 0;  // the wavefront's chunk [next, end) (the same in every lane)
  bool pool_empty = false, exhausted = false, has = false;
  uint32_t k = 0, prim = 0, c = 0;
  if (MODE == MODE_WRITE) {  // the write pass walks the listed leaves (usually none)
    n = counters[4];
    if (n == 0) return;
  }
  size_t w = 0;
  float q[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int sp = 0;
  uint32_t node = 0, top = 0;
  bool top_valid = false;
  unsigned long long nvis = 0;
  uint32_t dir = WQ_NODIR, chunk = 0;  // count pass: the leaf's directory and current chunk
  auto emit = [&](uint32_t leafpos) {
    if (MODE == MODE_WRITE) {
      // a pair past the buffered ones: its leaf position goes to pb[w], which the copy kernel,
      // run after this pass, maps to a shape and orders (no dependent gather in here)
      if (c >= WQ_LOCAL) {
        if (w < cap) pb[w] = leafpos;
        else counters[1] = 1u;
      }
      w++;
      c++;
    } else {
      if (c < WQ_LOCAL) {
        lbuf[(size_t)k * WQ_LOCAL + c] = leafpos;  // the emit kernel maps it to a shape
      } else {
        const uint32_t j = c - WQ_LOCAL;
        if (j == 0) {  // the first partner past the buffer: a directory
          dir = atomicAdd(&counters[2], 1u);
          if (dir >= ndir) dir = WQ_NODIR;
        }
        if (dir != WQ_NODIR && j % WQ_OCH == 0) {  // a new chunk
          if (j / WQ_OCH < WQ_DIRN) chunk = atomicAdd(&counters[3], 1u);
          if (j / WQ_OCH >= WQ_DIRN || chunk >= nch) dir = WQ_NODIR;
          else odirs[(size_t)dir * WQ_DIRN + j / WQ_OCH] = chunk;
        }
        if (dir != WQ_NODIR) ochunks[(size_t)chunk * WQ_OCH + j % WQ_OCH] = leafpos;
      }
      c++;
    }
  };
  for (;;) {
    // refill lanes without a leaf
    const bool want = !has && !exhausted;
    const uint64_t need = __ballot(want);
    if (need) {
      if (next >= end && !pool_empty) {
        const int leader = __ffsll((unsigned long long)need) - 1;
        while (part_k < WQ_PARTS) {
          const uint32_t p = (part0 + part_k) % WQ_PARTS;
          const uint32_t lo = (uint32_t)(((uint64_t)n * p) / WQ_PARTS);
          const uint32_t hi = (uint32_t)(((uint64_t)n * (p + 1)) / WQ_PARTS);
          uint32_t base = 0;
          if (lane == leader) base = atomicAdd(work + p * WQ_PART_STRIDE, WQ_CHUNK);
          base = __shfl(base, leader, 64);
          if (base < hi - lo) {
            next = lo + base;
            end = hi - next > WQ_CHUNK ? next + WQ_CHUNK : hi;
            break;
          }
          ++part_k;
        }
        if (part_k >= WQ_PARTS) pool_empty = true;
      }
      const uint32_t avail = end - next;
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
      if (want) {
        if (rank < avail) {
          k = MODE == MODE_WRITE ? relist[next + rank] : next + rank;
          prim = order[k];
          const float4 b0 = sbox[2 * (size_t)prim], b1 = sbox[2 * (size_t)prim + 1];
          q[0] = b0.x; q[1] = b0.y; q[2] = b0.z; q[3] = b0.w; q[4] = b1.x; q[5] = b1.y;
          c = 0;
          dir = WQ_NODIR;
          w = MODE == MODE_WRITE ? offs[k] : 0;
          sp = 0;
          node = 0;
          top_valid = false;
          has = true;
        } else if (pool_empty) {
          exhausted = true;
        }
      }
      const uint32_t taken = (uint32_t)__popcll(need);
      next += taken < avail ? taken : avail;
    }
    if (__ballot(has) == 0) {
      if (__ballot(!exhausted) == 0) break;
      continue;
    }
    if (!has) continue;
    nvis++;
    const float4* nq = (const float4*)(nodes + node);
    const float4 q0 = nq[0], q1 = nq[1], q2 = nq[2], q3 = nq[3];
    const uint32_t cl = __float_as_uint(q3.x), cr = __float_as_uint(q3.y);
    const uint32_t ll = __float_as_uint(q3.z), lr = __float_as_uint(q3.w);
    bool hl = ll > k && wbox_overlap(q, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y);
    bool hr = lr > k && wbox_overlap(q, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w);
    if (hl && (cl & NCX_LEAF_BIT)) { emit(cl & ~NCX_LEAF_BIT); hl = false; }
    if (hr && (cr & NCX_LEAF_BIT)) { emit(cr & ~NCX_LEAF_BIT); hr = false; }
    bool done = false;
    if (hl && hr) {
      if (top_valid) {
        if (sp < WQ_LSTACK) lstack[sp * WQ_BLOCK + tid] = top;
        else gs[sp - WQ_LSTACK] = top;
        sp++;
      }
      top = cr;
      top_valid = true;
      node = cl;
    } else if (hl) {
      node = cl;
    } else if (hr) {
      node = cr;
    } else if (top_valid) {
      node = top;
      top_valid = false;
    } else if (sp == 0) {
      done = true;
    } else {
      sp--;
      node = sp < WQ_LSTACK ? lstack[sp * WQ_BLOCK + tid] : gs[sp - WQ_LSTACK];
    }
    if (done) {
      if (MODE == MODE_COUNT) {
        cnt[k] = c;
        if (c > WQ_LOCAL) {
          ovdir[k] = dir;
          if (dir == WQ_NODIR) relist[atomicAdd(&counters[4], 1u)] = k;
        }
      }
      has = false;
    }
  }
  if (visits) atomicAdd(visits, nvis);  // internal nodes visited (measurement only)
}

// Persistent grid of the query kernels: about as many blocks as are resident, at most one lane
// per leaf.
static unsigned world_query_blocks(const ncx_ctx* c, size_t n) {
  const unsigned resident = (unsigned)c->num_cus * WQ_BLOCKS_PER_CU;
  return std::max(1u, std::min<unsigned>(resident, ceil_div(n, WQ_BLOCK)));
}

// Every pair, in the order the traversal found it, so the list is the same as one full write
// traversal per leaf would give: a leaf's first WQ_LOCAL partners from the count pass's buffer,
// the rest from its chunks (or, for a leaf marked WQ_NODIR, from the leaf positions the write pass
// left in pb); each is mapped to its shape and the pair ordered (a < b).  A block takes EMIT_LEAVES consecutive leaves; its
// threads walk the concatenation of those leaves' pairs (a block-wide prefix sum over the
// counts, then a binary search for each pair's leaf), so consecutive threads write consecutive
// pair slots.
#define WQ_EMIT_BLOCK 256
#define EMIT_LEAVES 64  // leaves per block: 256 leaves per block left dense clusters to few blocks
#define EMIT_UNROLL 2
static_assert(EMIT_LEAVES == 64, "the emit kernel's prefix sum over a block's leaf counts is one wavefront's scan");
static_assert(WQ_EMIT_BLOCK >= EMIT_LEAVES, "the first wavefront holds one leaf count per lane");
__global__ __launch_bounds__(WQ_EMIT_BLOCK) void k_world_emit(const uint32_t* __restrict__ order, uint32_t n,
                                                              const uint32_t* __restrict__ cnt,
                                                              const uint32_t* __restrict__ offs,
                                                              const uint32_t* __restrict__ lbuf, size_t cap,
                                                              uint32_t* __restrict__ pa, uint32_t* __restrict__ pb,
                                                              uint32_t* __restrict__ counters,
                                                              const uint32_t* __restrict__ ovdir,
                                                              const uint32_t* __restrict__ odirs,
                                                              const uint32_t* __restrict__ ochunks) {
  __shared__ uint32_t s_pre[EMIT_LEAVES + 1];
  const uint32_t t = threadIdx.x;
  const uint32_t k0 = blockIdx.x * EMIT_LEAVES;
  // every pair of the block's leaves: the first WQ_LOCAL of a leaf from the count pass's buffer,
  // the rest from its chunks or the leaf positions the write pass left in pb
  if (t < EMIT_LEAVES) {  // the first wavefront, all lanes active: inclusive prefix sum of the counts
    const uint32_t k = k0 + t;
    uint32_t x = k < n ? cnt[k] : 0u;
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
      if (lane >= d) x += y;
    }
    s_pre[t + 1] = x;
    if (t == 0) s_pre[0] = 0;
  }
  __syncthreads();
  const uint32_t total = s_pre[EMIT_LEAVES];
  // EMIT_UNROLL pairs per thread per pass, their loads issued together: a block over a dense
  // cluster walks tens of thousands of pairs, and one load round trip per pair was its bound.
  for (uint32_t j0 = t; j0 < total; j0 += EMIT_UNROLL * WQ_EMIT_BLOCK) {
    uint32_t kl[EMIT_UNROLL], ii[EMIT_UNROLL], w0[EMIT_UNROLL], prim[EMIT_UNROLL], h2[EMIT_UNROLL];
#pragma unroll
    for (int u = 0; u < EMIT_UNROLL; ++u) {
      // the leaf l with s_pre[l] <= j < s_pre[l + 1] (a pair past the end repeats the first one)
      const uint32_t j = j0 + u * WQ_EMIT_BLOCK < total ? j0 + u * WQ_EMIT_BLOCK : j0;
      uint32_t lo = 0, hi = EMIT_LEAVES;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (s_pre[mid] <= j) lo = mid;
        else hi = mid;
      }
      kl[u] = k0 + lo;
      ii[u] = j - s_pre[lo];
    }
#pragma unroll
    for (int u = 0; u < EMIT_UNROLL; ++u) {
      w0[u] = offs[kl[u]];
      prim[u] = order[kl[u]];
      const size_t w = (size_t)w0[u] + ii[u];
      uint32_t lp;
      if (ii[u] < WQ_LOCAL) {
        lp = lbuf[(size_t)kl[u] * WQ_LOCAL + ii[u]];
      } else {
        const uint32_t dd = ovdir[kl[u]], j = ii[u] - WQ_LOCAL;
        if (dd != WQ_NODIR) lp = ochunks[(size_t)odirs[(size_t)dd * WQ_DIRN + j / WQ_OCH] * WQ_OCH + j % WQ_OCH];
        else lp = w < cap ? pb[w] : 0u;
      }
      h2[u] = order[lp];
    }
#pragma unroll
    for (int u = 0; u < EMIT_UNROLL; ++u) {
      if (j0 + u * WQ_EMIT_BLOCK >= total) break;
      const size_t w = (size_t)w0[u] + ii[u];
      if (w < cap) {
        pa[w] = prim[u] < h2[u] ? prim[u] : h2[u];
        pb[w] = prim[u] < h2[u] ? h2[u] : prim[u];
      } else {
        counters[1] = 1u;
      }
    }
  }
}

// Enqueues the LBVH build (no host synchronisation, so a rebuild step can be captured).
static int world_build(ncx_world* w) {
  ncx_ctx* c = &w->wctx;
  const float* box[6];
  for (int d = 0; d < 6; ++d) box[d] = w->stored.as<float>() + d * w->n;
  int rc = lbvh_build(c, w->bvh, box, w->n, 0.f, w->sbox.as<float4>());
  if (rc) return rc;
  if (w->n > 1) {
    // The search stack holds fewer entries than the tree has levels (at most LBVH_MAX_DEPTH): one
    // in a register, WQ_LSTACK in LDS, the rest here.  A fixed bound needs no host read-back.
    w->spill = LBVH_MAX_DEPTH + 1 - WQ_LSTACK;
    if ((rc = w->gstack.reserve((size_t)world_query_blocks(c, w->n) * WQ_BLOCK * w->spill * 4))) return rc;
  }
  return NCX_OK;
}

// Enqueues steps 1, 2 and 4, 5 (and refit when `refit`) on the world stream.
static int world_enqueue(ncx_world* w, const ncx_pose* poses, bool refit) {
  ncx_ctx* c = &w->wctx;
  size_t n = w->n;
  int rc;
  // shape boxes and the stored-box rule in one kernel
  if ((rc = shapes_aabbs_stage_on(w->shapes, c, poses, w->prediction, w->stored.as<float>(), w->sbox.as<float4>(),
                                  w->margin, w->first)))
    return rc;
  if (refit) {
    const float* box[6];
    for (int d = 0; d < 6; ++d) box[d] = w->stored.as<float>() + d * n;
    if ((rc = lbvh_refit(c, w->bvh, box, 0.f, w->sbox.as<float4>()))) return rc;
  }
  return NCX_OK;
}

static int world_pairs_and_contacts(ncx_world* w, const ncx_pose* poses, bool eager) {
  ncx_ctx* c = &w->wctx;
  size_t n = w->n;
  int rc;
  uint32_t* ctr = w->counters.as<uint32_t>();
  const unsigned tw_blocks = std::max(1u, std::min<unsigned>(ceil_div(w->cap, 256), (unsigned)c->num_cus * 8));
  if (w->temporal && !w->first) {
    // the previous step's pairs (still in pa / pb, counted in ctr[0]) and result directions
    if ((rc = fill_u32(c, w->tw_keys.as<uint32_t>(), 2 * (w->tw_mask + 1), 0xFFFFFFFFu))) return rc;
    hipLaunchKernelGGL(k_tw_insert, dim3(tw_blocks), dim3(256), 0, c->stream, w->pa.as<uint32_t>(),
                       w->pb.as<uint32_t>(), ctr, w->cap, w->shapes->shapes.as<ncx_shape>(), w->tw_fdir.as<float>(),
                       w->tw_keys.as<unsigned long long>(), w->tw_dirs.as<float4>(), w->tw_mask);
    NCX_HIP_CHECK(hipGetLastError());
  }
  if ((rc = fill_u32(c, ctr, 5, 0u))) return rc;
  if (n >= 2) {
    unsigned nb = world_query_blocks(c, n);
    const LNode* nodes = w->bvh.nodes.as<LNode>();
    const uint32_t* order = w->bvh.order.as<uint32_t>();
    const float4* st = w->sbox.as<float4>();
    uint32_t* qw = w->qwork.as<uint32_t>();
    if ((rc = fill_u32(c, qw, 2 * WQ_PARTS * WQ_PART_STRIDE, 0u))) return rc;
    hipLaunchKernelGGL((k_world_query<MODE_COUNT>), dim3(nb), dim3(WQ_BLOCK), 0, c->stream, nodes, order, st,
                       (uint32_t)n, w->cnt.as<uint32_t>(), (const uint32_t*)nullptr, w->cap, (uint32_t*)nullptr,
                       (uint32_t*)nullptr, ctr, w->gstack.as<uint32_t>(), w->spill, w->lbuf.as<uint32_t>(),
                       (unsigned long long*)nullptr, qw, w->ovdir.as<uint32_t>(), w->odirs.as<uint32_t>(), w->ndir,
                       w->ochunks.as<uint32_t>(), w->nch, w->relist.as<uint32_t>());
    NCX_HIP_CHECK(hipGetLastError());
    if ((rc = scan_exclusive_u32(c, w->cnt.as<uint32_t>(), w->offs.as<uint32_t>(), n, w->bvh.scan_tmp, ctr)))
      return rc;
    // the write pass (leaves marked WQ_NODIR only) first: it leaves leaf positions in pb for the
    // copy kernel to map
    hipLaunchKernelGGL((k_world_query<MODE_WRITE>), dim3(nb), dim3(WQ_BLOCK), 0, c->stream, nodes, order, st,
                       (uint32_t)n, w->cnt.as<uint32_t>(), w->offs.as<uint32_t>(), w->cap, w->pa.as<uint32_t>(),
                       w->pb.as<uint32_t>(), ctr, w->gstack.as<uint32_t>(), w->spill, w->lbuf.as<uint32_t>(),
                       (unsigned long long*)nullptr, qw + WQ_PARTS * WQ_PART_STRIDE, w->ovdir.as<uint32_t>(),
                       w->odirs.as<uint32_t>(), w->ndir, w->ochunks.as<uint32_t>(), w->nch, w->relist.as<uint32_t>());
    NCX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_world_emit, dim3(ceil_div(n, EMIT_LEAVES)), dim3(WQ_EMIT_BLOCK), 0, c->stream, order, (uint32_t)n,
                       w->cnt.as<uint32_t>(), w->offs.as<uint32_t>(), w->lbuf.as<uint32_t>(), w->cap,
                       w->pa.as<uint32_t>(), w->pb.as<uint32_t>(), ctr, w->ovdir.as<uint32_t>(),
                       w->odirs.as<uint32_t>(), w->ochunks.as<uint32_t>());
    NCX_HIP_CHECK(hipGetLastError());
  }
  if (!w->contacts) return NCX_OK;
  const float* idir = nullptr;
  if (w->temporal) {
    if (w->first) {
      if ((rc = fill_u32(c, w->tw_init.as<uint32_t>(), 3 * w->cap, 0u))) return rc;
    } else {
      hipLaunchKernelGGL(k_tw_lookup, dim3(tw_blocks), dim3(256), 0, c->stream, w->pa.as<uint32_t>(),
                         w->pb.as<uint32_t>(), ctr, w->cap, w->shapes->shapes.as<ncx_shape>(),
                         w->tw_keys.as<unsigned long long>(), w->tw_dirs.as<float4>(), w->tw_mask,
                         w->tw_init.as<float>());
      NCX_HIP_CHECK(hipGetLastError());
    }
    return contacts_launch_device_count(w->shapes, w->tiers, c, poses, w->pa.as<uint32_t>(), w->pb.as<uint32_t>(),
                                        ctr, w->cap, w->prediction, w->tw_init.as<float>(), 10000,
                                        w->has.as<uint8_t>(), w->w1.as<float>(), w->w2.as<float>(),
                                        w->nrm.as<float>(), w->depth_out.as<float>(), w->iters.as<int32_t>(), eager,
                                        w->tw_fdir.as<float>());
  }
  if (w->mlp) {
    if ((rc = mlp_predict_device_count(w->mlp, c, w->shapes->shapes.as<ncx_shape>(), poses, w->desc.as<float>(),
                                       w->pa.as<uint32_t>(), w->pb.as<uint32_t>(), ctr, w->cap, w->dirs.as<float>())))
      return rc;
    idir = w->dirs.as<float>();
  }
  // Pairs whose warm start hits the 32-iteration cap are recomputed by plain GJK inside the kernel.
  return contacts_launch_device_count(w->shapes, w->tiers, c, poses, w->pa.as<uint32_t>(), w->pb.as<uint32_t>(), ctr, w->cap,
                                      w->prediction, idir, w->mlp ? 32 : 10000, w->has.as<uint8_t>(),
                                      w->w1.as<float>(), w->w2.as<float>(), w->nrm.as<float>(),
                                      w->depth_out.as<float>(), w->iters.as<int32_t>(), eager);
}

extern "C" int ncx_world_create(ncx_ctx* ctx, ncx_shapes* shapes, float margin, float prediction, size_t max_pairs,
                                ncx_world** out) {
  NCX_REQUIRE(ctx && shapes && out, "ncx_world_create: null argument");
  NCX_REQUIRE(margin >= 0.f, "ncx_world_create: the loosening margin must be positive");
  NCX_REQUIRE(max_pairs > 0 && max_pairs < 0xFFFFFFFFull, "ncx_world_create: max_pairs out of range");
  ncx_world* w = new ncx_world;
  w->ctx = ctx;
  w->wctx = *ctx;
  w->shapes = shapes;
  w->n = shapes->n;
  w->cap = max_pairs;
  w->margin = margin;
  w->prediction = prediction;
  size_t n = w->n, cap = max_pairs;
  // Directories for one leaf in 64 (at least 64, at most 65,536); the chunk pool holds every pair
  // the pair list can hold plus one partly filled chunk per directory.
  w->ndir = (uint32_t)std::min<size_t>(std::max<size_t>(n / 64, 64), 65536);
  w->nch = (uint32_t)std::min<size_t>(cap / WQ_OCH + w->ndir, 0xFFFFFFFFull / WQ_OCH);
  int rc;
  if (hipStreamCreateWithFlags(&w->wctx.stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&w->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&w->ev_out, hipEventDisableTiming) != hipSuccess) {
    delete w;
    ncx_set_error("ncx_world_create: stream/event creation failed");
    return NCX_E_HIP;
  }
  if ((rc = w->stored.reserve(n * 24 + 64)) ||
      (rc = w->sbox.reserve(n * 32 + 64)) ||
      (rc = w->cnt.reserve(n * 4 + 64)) || (rc = w->offs.reserve(n * 4 + 64)) || (rc = w->counters.reserve(64)) ||
      (rc = w->lbuf.reserve(n * (size_t)WQ_LOCAL * 4 + 64)) || (rc = w->ovdir.reserve(n * 4 + 64)) ||
      (rc = w->relist.reserve(n * 4 + 64)) ||
      (rc = w->qwork.reserve(2 * WQ_PARTS * WQ_PART_STRIDE * 4)) ||
      (rc = w->odirs.reserve((size_t)w->ndir * WQ_DIRN * 4 + 64)) ||
      (rc = w->ochunks.reserve((size_t)w->nch * WQ_OCH * 4 + 64)) ||
      (rc = w->pa.reserve(cap * 4)) || (rc = w->pb.reserve(cap * 4)) || (rc = w->has.reserve(cap)) ||
      (rc = w->w1.reserve(cap * 12)) || (rc = w->w2.reserve(cap * 12)) || (rc = w->nrm.reserve(cap * 12)) ||
      (rc = w->depth_out.reserve(cap * 4)) || (rc = w->iters.reserve(cap * 4)) || (rc = contact_tiers_reserve(w->tiers, std::min<size_t>(cap, (size_t)1 << 22)))) {
    contact_tiers_release(w->tiers);
    delete w;
    return rc;
  }
  *out = w;
  return NCX_OK;
}

extern "C" int ncx_world_set_warmstart(ncx_world* w, ncx_mlp* mlp) {
  NCX_REQUIRE(w, "ncx_world_set_warmstart: null world");
  NCX_REQUIRE(!(mlp && w->temporal), "ncx_world_set_warmstart: the temporal warm start is on");
  w->mlp = mlp;
  w->graph_valid = false;
  w->rb_valid = false;
  if (mlp) {
    int rc;
    if ((rc = w->desc.reserve(w->n * 16 * 4 + 64)) || (rc = w->dirs.reserve(w->cap * 12))) return rc;
    if ((rc = shape_descriptors_on(w->shapes, &w->wctx, w->desc.as<float>()))) return rc;
    NCX_HIP_CHECK(hipStreamSynchronize(w->wctx.stream));
  }
  return NCX_OK;
}

extern "C" int ncx_world_set_rebuild_every(ncx_world* w, int k) {
  NCX_REQUIRE(w && k >= 1, "ncx_world_set_rebuild_every: bad argument");
  w->rebuild_every = k;
  return NCX_OK;
}

// Turns the narrow phase of each step on or off (off: the step ends with the pair list).
extern "C" int ncx_world_set_contacts(ncx_world* w, int enable) {
  NCX_REQUIRE(w, "ncx_world_set_contacts: null argument");
  if (w->contacts != (enable != 0)) {
    w->graph_valid = false;
    w->rb_valid = false;
  }
  w->contacts = enable != 0;
  return NCX_OK;
}

// Measurement: internal nodes visited by one pair-count traversal of the current tree and
// stored boxes (the work of each of the step's two query passes).  Synchronous.
extern "C" int ncx_world_query_stats(ncx_world* w, unsigned long long* nodes_visited) {
  NCX_REQUIRE(w && nodes_visited, "ncx_world_query_stats: null argument");
  NCX_REQUIRE(!w->first, "ncx_world_query_stats: no step has run yet");
  *nodes_visited = 0;
  if (w->n < 2) return NCX_OK;
  ncx_ctx* c = &w->wctx;
  DevBuf v;
  int rc;
  if ((rc = v.reserve(8))) return rc;
  if ((rc = fill_u32(c, v.as<uint32_t>(), 2, 0u))) return rc;
  if ((rc = fill_u32(c, w->counters.as<uint32_t>() + 2, 3, 0u))) return rc;
  unsigned nb = world_query_blocks(c, w->n);
  if ((rc = fill_u32(c, w->qwork.as<uint32_t>(), WQ_PARTS * WQ_PART_STRIDE, 0u))) return rc;
  hipLaunchKernelGGL((k_world_query<MODE_COUNT>), dim3(nb), dim3(WQ_BLOCK), 0, c->stream, w->bvh.nodes.as<LNode>(),
                     w->bvh.order.as<uint32_t>(), w->sbox.as<float4>(), (uint32_t)w->n, w->cnt.as<uint32_t>(),
                     (const uint32_t*)nullptr, w->cap, (uint32_t*)nullptr, (uint32_t*)nullptr,
                     w->counters.as<uint32_t>(), w->gstack.as<uint32_t>(), w->spill, w->lbuf.as<uint32_t>(),
                     v.as<unsigned long long>(), w->qwork.as<uint32_t>(), w->ovdir.as<uint32_t>(),
                     w->odirs.as<uint32_t>(), w->ndir, w->ochunks.as<uint32_t>(), w->nch, w->relist.as<uint32_t>());
  NCX_HIP_CHECK(hipGetLastError());
  NCX_HIP_CHECK(hipMemcpyAsync(nodes_visited, v.p, 8, hipMemcpyDeviceToHost, c->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(c->stream));
  v.release();
  return NCX_OK;
}

extern "C" int ncx_world_spill_stats(ncx_world* w, uint32_t out[2]) {
  NCX_REQUIRE(w && out, "ncx_world_spill_stats: null argument");
  out[0] = out[1] = 0;
  if (w->first || w->n < 2) return NCX_OK;
  ncx_ctx* c = &w->wctx;
  std::vector<uint32_t> cnt(w->n), dir(w->n), ctr(4);
  NCX_HIP_CHECK(hipMemcpyAsync(cnt.data(), w->cnt.p, w->n * 4, hipMemcpyDeviceToHost, c->stream));
  NCX_HIP_CHECK(hipMemcpyAsync(dir.data(), w->ovdir.p, w->n * 4, hipMemcpyDeviceToHost, c->stream));
  NCX_HIP_CHECK(hipMemcpyAsync(ctr.data(), w->counters.p, 16, hipMemcpyDeviceToHost, c->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(c->stream));
  out[0] = ctr[2];
  for (size_t k = 0; k < w->n; ++k)
    if (cnt[k] > WQ_LOCAL && dir[k] != WQ_NODIR) out[1]++;
  return NCX_OK;
}

extern "C" int ncx_world_destroy(ncx_world* w) {
  if (!w) return NCX_OK;
  if (w->exec) (void)hipGraphExecDestroy(w->exec);
  if (w->graph) (void)hipGraphDestroy(w->graph);
  if (w->exec_rb) (void)hipGraphExecDestroy(w->exec_rb);
  if (w->graph_rb) (void)hipGraphDestroy(w->graph_rb);
  if (w->wctx.stream) {
    (void)hipStreamSynchronize(w->wctx.stream);
    (void)hipStreamDestroy(w->wctx.stream);
  }
  if (w->ev_in) (void)hipEventDestroy(w->ev_in);
  if (w->ev_out) (void)hipEventDestroy(w->ev_out);
  DevBuf* bufs[] = {&w->stored, &w->sbox, &w->gstack, &w->qwork, &w->lbuf, &w->ovdir, &w->odirs,
                    &w->ochunks, &w->relist, &w->cnt, &w->offs, &w->counters, &w->pa, &w->pb,
                    &w->has, &w->w1, &w->w2, &w->nrm, &w->depth_out, &w->iters, &w->desc, &w->dirs,
                    &w->qcnt, &w->qscan, &w->qtot, &w->tw_keys, &w->tw_dirs, &w->tw_init, &w->tw_fdir,
                    &w->rkeys, &w->rvals, &w->rtk, &w->rtv, &w->rhist, &w->rscan};
  for (DevBuf* b : bufs) b->release();
  w->bvh.release();
  contact_tiers_release(w->tiers);
  delete w;
  return NCX_OK;
}

// Captures the step's work (enqueue) into *exec on stream s.
template <class F>
static int capture_step(hipStream_t s, hipGraph_t* graph, hipGraphExec_t* exec, F enqueue) {
  if (*exec) { (void)hipGraphExecDestroy(*exec); *exec = nullptr; }
  if (*graph) { (void)hipGraphDestroy(*graph); *graph = nullptr; }
  NCX_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int r = enqueue();
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (r) { if (g) (void)hipGraphDestroy(g); return r; }
  NCX_HIP_CHECK(e);
  *graph = g;
  NCX_HIP_CHECK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return NCX_OK;
}

// One step.  use_graph != 0: refit steps and rebuild steps (after the first step, which tunes
// and allocates) each replay a captured hipGraph, captured on first use and again when
// `poses_dev` changes.  Neither depends on the tree's topology, so a rebuild does not invalidate
// the refit graph.  The step is ordered after prior work on the caller's stream and later caller
// work waits for it; ncx_world_sync waits for completion.
extern "C" int ncx_world_sync(ncx_world* w, size_t* npairs, size_t* n_epa_overflow);

extern "C" int ncx_world_step(ncx_world* w, const ncx_pose* poses_dev, int use_graph) {
  NCX_REQUIRE(w && poses_dev, "ncx_world_step: null argument");
  if (w->temporal && !w->first && !w->synced) {
    // every result direction of the last step (including pairs the host-driven EPA tier finishes
    // at sync) must be in place before this step reads them
    size_t np = 0;
    int r = ncx_world_sync(w, &np, nullptr);
    if (r) return r;
  }
  w->synced = false;
  hipStream_t s = w->wctx.stream;
  NCX_HIP_CHECK(hipEventRecord(w->ev_in, w->ctx->stream));
  NCX_HIP_CHECK(hipStreamWaitEvent(s, w->ev_in, 0));
  int rc;
  bool rebuild = w->first || w->since_build >= w->rebuild_every;
  w->last_was_graph = 0;
  if (rebuild && (w->first || !use_graph)) {
    if ((rc = world_enqueue(w, poses_dev, false))) return rc;
    if ((rc = world_build(w))) return rc;
    if ((rc = world_pairs_and_contacts(w, poses_dev, true))) return rc;
  } else if (rebuild) {
    if (!w->rb_valid || w->rb_poses != poses_dev) {
      rc = capture_step(s, &w->graph_rb, &w->exec_rb, [&]() {
        int r = world_enqueue(w, poses_dev, false);
        if (!r) r = world_build(w);
        if (!r) r = world_pairs_and_contacts(w, poses_dev, false);
        return r;
      });
      if (rc) return rc;
      w->rb_valid = true;
      w->rb_poses = poses_dev;
    }
    NCX_HIP_CHECK(hipGraphLaunch(w->exec_rb, s));
    w->last_was_graph = 1;
  } else if (use_graph) {
    if (!w->graph_valid || w->graph_poses != poses_dev) {
      rc = capture_step(s, &w->graph, &w->exec, [&]() {
        int r = world_enqueue(w, poses_dev, true);
        if (!r) r = world_pairs_and_contacts(w, poses_dev, false);
        return r;
      });
      if (rc) return rc;
      w->graph_valid = true;
      w->graph_poses = poses_dev;
    }
    NCX_HIP_CHECK(hipGraphLaunch(w->exec, s));
    w->last_was_graph = 1;
  } else {
    if ((rc = world_enqueue(w, poses_dev, true))) return rc;
    if ((rc = world_pairs_and_contacts(w, poses_dev, true))) return rc;
  }
  if (rebuild) w->since_build = 0;
  w->first = false;
  w->since_build++;
  w->last_had_contacts = w->contacts;
  NCX_HIP_CHECK(hipEventRecord(w->ev_out, s));
  NCX_HIP_CHECK(hipStreamWaitEvent(w->ctx->stream, w->ev_out, 0));
  return NCX_OK;
}

// Waits for the last step and reports its pair count.  Fails on a pair-capacity overflow.
// Contacts whose EPA polytope outgrew both in-step arena tiers are recomputed here with larger
// arenas (the step's poses must be unchanged until this returns); *n_epa_overflow reports how
// many pairs of the step needed more than the first-tier arena.
extern "C" int ncx_world_sync(ncx_world* w, size_t* npairs, size_t* n_epa_overflow) {
  NCX_REQUIRE(w && npairs, "ncx_world_sync: null argument");
  uint32_t ctr[2] = {0, 0};
  NCX_HIP_CHECK(hipMemcpyAsync(ctr, w->counters.p, 8, hipMemcpyDeviceToHost, w->wctx.stream));
  NCX_HIP_CHECK(hipStreamSynchronize(w->wctx.stream));
  *npairs = ctr[0];
  if (ctr[1] || ctr[0] > w->cap) {
    ncx_set_error("ncx_world_sync: more pairs than max_pairs");
    return NCX_E_CAPACITY;
  }
  size_t large = 0;
  if (w->last_had_contacts) {
    int rc = contacts_resolve_overflow(w->tiers, &w->wctx, &large);
    if (rc) return rc;
  }
  if (n_epa_overflow) *n_epa_overflow = large;
  w->synced = true;
  return NCX_OK;
}

extern "C" int ncx_world_outputs(ncx_world* w, uint32_t** ia, uint32_t** ib, uint8_t** has, float** world1,
                                 float** world2, float** normal, float** depth, int32_t** gjk_iters) {
  NCX_REQUIRE(w, "ncx_world_outputs: null argument");
  if (ia) *ia = w->pa.as<uint32_t>();
  if (ib) *ib = w->pb.as<uint32_t>();
  if (has) *has = w->has.as<uint8_t>();
  if (world1) *world1 = w->w1.as<float>();
  if (world2) *world2 = w->w2.as<float>();
  if (normal) *normal = w->nrm.as<float>();
  if (depth) *depth = w->depth_out.as<float>();
  if (gjk_iters) *gjk_iters = w->iters.as<int32_t>();
  return NCX_OK;
}

// The stored (fat) box of every proxy after the last step, device SoA (6 x n floats).
extern "C" int ncx_world_set_temporal_warmstart(ncx_world* w, int enable) {
  NCX_REQUIRE(w, "ncx_world_set_temporal_warmstart: null argument");
  NCX_REQUIRE(!(enable && w->mlp), "ncx_world_set_temporal_warmstart: the learned warm start is set");
  if (enable && !w->temporal) {
    uint64_t slots = 1;
    while (slots < 2 * (uint64_t)w->cap) slots <<= 1;
    int rc;
    if ((rc = w->tw_keys.reserve(slots * 8)) || (rc = w->tw_dirs.reserve(slots * 16)) ||
        (rc = w->tw_init.reserve(w->cap * 12 + 16)) || (rc = w->tw_fdir.reserve(w->cap * 12 + 16)))
      return rc;
    w->tw_mask = slots - 1;
    // no pair has a direction yet: the first temporal step inserts nothing
    if ((rc = fill_u32(&w->wctx, w->tw_fdir.as<uint32_t>(), 3 * w->cap, 0u))) return rc;
    NCX_HIP_CHECK(hipStreamSynchronize(w->wctx.stream));
  }
  if ((enable != 0) != w->temporal) {
    w->temporal = enable != 0;
    w->graph_valid = false;  // the captured steps hold the other narrow phase
    w->rb_valid = false;
  }
  return NCX_OK;
}

extern "C" int ncx_world_temporal_dirs(ncx_world* w, const float** init_dev, const float** result_dev) {
  NCX_REQUIRE(w && init_dev && result_dev, "ncx_world_temporal_dirs: null argument");
  NCX_REQUIRE(w->temporal, "ncx_world_temporal_dirs: the temporal warm start is off");
  *init_dev = w->tw_init.as<float>();
  *result_dev = w->tw_fdir.as<float>();
  return NCX_OK;
}

extern "C" int ncx_world_warmstart_dirs(ncx_world* w, const float** dirs_dev) {
  NCX_REQUIRE(w && dirs_dev, "ncx_world_warmstart_dirs: null argument");
  NCX_REQUIRE(w->mlp, "ncx_world_warmstart_dirs: no learned warm start set");
  *dirs_dev = w->dirs.as<float>();
  return NCX_OK;
}

extern "C" int ncx_world_stored_aabbs(ncx_world* w, const float** soa) {
  NCX_REQUIRE(w && soa, "ncx_world_stored_aabbs: null argument");
  *soa = w->stored.as<float>();
  return NCX_OK;
}

// ------------------------------------------------------------------ interference queries of a scene
// On the caller's stream, after the last step (which that stream waits for): the tree and the
// stored boxes are those of the last step, the shapes are at poses_dev.
static BQArgs world_bq_args(ncx_world* w, const ncx_pose* poses) {
  BQArgs a = {};
  a.nodes = w->bvh.nodes.as<LNode>();
  a.order = w->bvh.order.as<uint32_t>();
  a.alive = nullptr;
  a.in_tree = nullptr;
  a.leaf_soa = w->stored.as<float>();
  a.nleaves = (uint32_t)w->n;
  a.shapes = w->shapes->shapes.as<ncx_shape>();
  a.P = Pool{w->shapes->hx.as<float>(), w->shapes->hy.as<float>(), w->shapes->hz.as<float>()};
  a.poses = poses;
  return a;
}

extern "C" int ncx_world_interferences_with_rays(ncx_world* w, const ncx_pose* poses_dev,
                                                 const float* const ray_soa_dev[6], size_t n, float max_toi,
                                                 int aabb_only, uint32_t* offsets_dev, uint32_t* handles_dev,
                                                 float* toi_dev, float* normal_dev, size_t cap, size_t* total) {
  NCX_REQUIRE(w && ray_soa_dev && offsets_dev && handles_dev && total && (aabb_only || poses_dev),
              "ncx_world_interferences_with_rays: null argument");
  NCX_REQUIRE(!w->first, "ncx_world_interferences_with_rays: no step has run");
  NCX_REQUIRE(n < 0xFFFFFFFFull, "ncx_world_interferences_with_rays: too many rays");
  for (int d = 0; d < 6; ++d) NCX_REQUIRE(ray_soa_dev[d], "ncx_world_interferences_with_rays: null ray array");
  BQArgs a = world_bq_args(w, poses_dev);
  for (int d = 0; d < 6; ++d) a.q[d] = ray_soa_dev[d];
  a.nq = (uint32_t)n;
  a.max_toi = max_toi;
  a.cast_shapes = aabb_only ? 0 : 1;
  a.handles = handles_dev;
  a.toi = toi_dev;
  a.nrm = normal_dev;
  return bq_collect_run(w->ctx, a, BQ_RAY, offsets_dev, cap, total, w->qcnt, w->qscan, w->qtot);
}

extern "C" int ncx_world_interferences_with_aabbs(ncx_world* w, const float* const aabb_soa_dev[6], size_t n,
                                                  uint32_t* offsets_dev, uint32_t* handles_dev, size_t cap,
                                                  size_t* total) {
  NCX_REQUIRE(w && aabb_soa_dev && offsets_dev && handles_dev && total,
              "ncx_world_interferences_with_aabbs: null argument");
  NCX_REQUIRE(!w->first, "ncx_world_interferences_with_aabbs: no step has run");
  NCX_REQUIRE(n < 0xFFFFFFFFull, "ncx_world_interferences_with_aabbs: too many boxes");
  for (int d = 0; d < 6; ++d) NCX_REQUIRE(aabb_soa_dev[d], "ncx_world_interferences_with_aabbs: null box array");
  BQArgs a = world_bq_args(w, nullptr);
  for (int d = 0; d < 6; ++d) a.q[d] = aabb_soa_dev[d];
  a.nq = (uint32_t)n;
  a.handles = handles_dev;
  return bq_collect_run(w->ctx, a, BQ_BOX, offsets_dev, cap, total, w->qcnt, w->qscan, w->qtot);
}

// Sort key of a first-interference ray: the octant of its direction (bits 27-29), then the
// Morton code of its origin on a 512^3 grid over the tree's root box, clamped (bits 0-26).  Taken
// in key order, a wavefront's rays start close by and head the same way, so their traversals
// and leaf tests are alike.
struct RayKeyArgs {
  const float* q[6];
  const LNode* nodes;
  uint32_t n;
  uint32_t* keys;
  uint32_t* vals;
};

__device__ __forceinline__ uint32_t ray_spread10(uint32_t v) {  // bit k of v -> bit 3k
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ __launch_bounds__(256) void k_ray_keys(RayKeyArgs a) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= a.n) return;
  const LNode r = a.nodes[0];
  const float lo[3] = {fminf(r.lminx, r.rminx), fminf(r.lminy, r.rminy), fminf(r.lminz, r.rminz)};
  const float hi[3] = {fmaxf(r.lmaxx, r.rmaxx), fmaxf(r.lmaxy, r.rmaxy), fmaxf(r.lmaxz, r.rmaxz)};
  uint32_t key = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float ext = hi[d] - lo[d];
    const float u = ext > 0.f ? (a.q[d][i] - lo[d]) / ext : 0.f;
    const uint32_t c = (uint32_t)fminf(fmaxf(u * 512.f, 0.f), 511.f);  // NaN -> 0
    key |= ray_spread10(c) << (2 - d);
    if (a.q[3 + d][i] < 0.f) key |= 1u << (27 + d);
  }
  a.keys[i] = key;
  a.vals[i] = i;
}

// Batches of at least this many first-interference rays are taken in key order.
#define BQ_SORT_MIN 1024

extern "C" int ncx_world_first_interference_with_rays(ncx_world* w, const ncx_pose* poses_dev,
                                                      const float* const ray_soa_dev[6], size_t n, float max_toi,
                                                      uint32_t* handle_dev, float* toi_dev, float* normal_dev) {
  NCX_REQUIRE(w && poses_dev && ray_soa_dev && handle_dev && toi_dev && normal_dev,
              "ncx_world_first_interference_with_rays: null argument");
  NCX_REQUIRE(!w->first, "ncx_world_first_interference_with_rays: no step has run");
  NCX_REQUIRE(n < 0xFFFFFFFFull, "ncx_world_first_interference_with_rays: too many rays");
  for (int d = 0; d < 6; ++d) NCX_REQUIRE(ray_soa_dev[d], "ncx_world_first_interference_with_rays: null ray array");
  if (n == 0) return NCX_OK;
  BQArgs a = world_bq_args(w, poses_dev);
  for (int d = 0; d < 6; ++d) a.q[d] = ray_soa_dev[d];
  a.nq = (uint32_t)n;
  a.max_toi = max_toi;
  a.handles = handle_dev;
  a.toi = toi_dev;
  a.nrm = normal_dev;
  int rc;
  if ((rc = w->qtot.reserve(16)) || (rc = fill_u32(w->ctx, w->qtot.as<uint32_t>() + 1, 1, 0u))) return rc;
  a.err = w->qtot.as<uint32_t>() + 1;
  a.perm = nullptr;
  if (n >= BQ_SORT_MIN && w->n > 1) {
    if ((rc = w->rkeys.reserve(n * 4)) || (rc = w->rvals.reserve(n * 4))) return rc;
    RayKeyArgs k;
    for (int d = 0; d < 6; ++d) k.q[d] = ray_soa_dev[d];
    k.nodes = w->bvh.nodes.as<LNode>();
    k.n = (uint32_t)n;
    k.keys = w->rkeys.as<uint32_t>();
    k.vals = w->rvals.as<uint32_t>();
    hipLaunchKernelGGL(k_ray_keys, dim3(ceil_div(n, 256)), dim3(256), 0, w->ctx->stream, k);
    NCX_HIP_CHECK(hipGetLastError());
    if ((rc = radix_sort_u32(w->ctx, k.keys, k.vals, n, 0, 30, w->rtk, w->rtv, w->rhist, w->rscan))) return rc;
    a.perm = k.vals;  // a permutation of 0..n-1
  }
  hipLaunchKernelGGL(k_bq_first_ray<BQ_RAY_G>, dim3(ceil_div(n * BQ_RAY_G, BQ_BLOCK)), dim3(BQ_BLOCK), 0,
                     w->ctx->stream, a);
  NCX_HIP_CHECK(hipGetLastError());
  uint32_t err = 0;
  NCX_HIP_CHECK(hipMemcpyAsync(&err, w->qtot.as<uint32_t>() + 1, 4, hipMemcpyDeviceToHost, w->ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(w->ctx->stream));
  NCX_REQUIRE(err == 0, "ncx_world_first_interference_with_rays: traversal stack overflow");
  return NCX_OK;
}
