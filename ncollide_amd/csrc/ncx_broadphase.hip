// ncx_broadphase.hip -- DBVTBroadPhase replacement on the GPU.
//
// Reference: src/pipeline/broad_phase/dbvt_broad_phase.rs.  Proxies live in a
// host-side slab with the slab crate's LIFO key reuse, so handles are the ones
// the reference would hand out.  Stored boxes live on the device, one slot per
// handle, and follow the reference's rules: create_proxy stores the box as given
// (:275-280); deferred_set_bounding_volume stores box.loosened(margin) when the
// box is not contained in the stored one (:325-347).  update() recomputes the
// whole pair set {(a,b): stored_a intersects stored_b} with an LBVH self-query
// (each leaf only visits leaves after it in Morton order, so every pair is
// produced once), instead of re-inserting leaves into a DBVT.  The set equals
// the reference's HashMap contents after update (:174-259): intersection tests
// are exact f32 comparisons and node boxes are exact min/max unions.
#include "ncx_bvh_query.h"
#include <algorithm>

struct ncx_bp {
  ncx_ctx* ctx;
  float margin;
  // host slab (slab crate semantics: vacant entries form a LIFO free list)
  std::vector<uint8_t> occupied;
  std::vector<uint32_t> next_free;
  uint32_t slab_next = 0;
  size_t slab_len = 0;
  size_t cap = 0;            // device slots
  DevBuf stored;             // 6 * cap floats (SoA by component)
  DevBuf pending;            // 6 * cap floats
  DevBuf pending_flag;       // cap uint32: 1 = create (store as given), 2 = set (store loosened)
  DevBuf in_tree;            // cap uint32: 1 once the proxy went through an update
  // alive proxies compacted for the LBVH
  bool topology_dirty = true;
  std::vector<uint32_t> alive_host;
  DevBuf alive;              // uint32 handles
  DevBuf boxes;              // 6 * n_alive compacted boxes
  LBVH bvh;
  // pair lists (current and previous), as (a, b) with a < b
  DevBuf cnt, offs, total;
  DevBuf pa, pb, prev_a, prev_b;
  size_t npairs = 0, prev_npairs = 0;
  bool sorted = false, prev_sorted = false;
  DevBuf sort_t1, sort_t2, sort_t3, sort_hist, sort_scan;
  DevBuf ev_flags, ev_offs;
  DevBuf staging_h, staging_b;
  bool have_prev = false;
  // pairs dropped by remove() since the last update, reported there (the reference's remove
  // handler); they are taken out of the previous pair set when the next update's events are made
  std::vector<uint64_t> removed_pending, removed_prev;
  DevBuf qcnt, qscan, qtot;  // interference queries
  // collision groups per handle (membership, whitelist, blacklist), the pair filter; off until set
  bool groups_on = false, groups_dirty = false;
  std::vector<uint32_t> groups_host;
  DevBuf groups;
};

// CollisionGroups::new() (src/pipeline/object/collision_groups.rs:54-60): member of and
// interacting with all 30 groups.
#define NCX_ALL_GROUPS ((1u << 30) - 1u)

// CollisionGroups::can_interact_with_groups (collision_groups.rs:353-359).
__device__ __forceinline__ bool groups_allow(const uint32_t* __restrict__ g, uint32_t a, uint32_t b) {
  if (!g) return true;
  const uint32_t ma = g[3 * a], wa = g[3 * a + 1], ba = g[3 * a + 2];
  const uint32_t mb = g[3 * b], wb = g[3 * b + 1], bb = g[3 * b + 2];
  return (ma & bb) == 0 && (mb & ba) == 0 && (ma & wb) != 0 && (mb & wa) != 0;
}

static int bp_grow(ncx_bp* bp, size_t need) {
  if (need <= bp->cap) return NCX_OK;
  size_t ncap = bp->cap ? bp->cap : 1024;
  while (ncap < need) ncap *= 2;
  ncx_ctx* ctx = bp->ctx;
  DevBuf ns, np, nf, nt;
  int rc;
  if ((rc = ns.reserve(ncap * 24)) || (rc = np.reserve(ncap * 24)) || (rc = nf.reserve(ncap * 4)) ||
      (rc = nt.reserve(ncap * 4))) {
    ns.release(); np.release(); nf.release(); nt.release();
    return rc;
  }
  if ((rc = fill_u32(ctx, nf.as<uint32_t>(), ncap, 0u)) || (rc = fill_u32(ctx, nt.as<uint32_t>(), ncap, 0u))) return rc;
  for (int d = 0; d < 6; ++d) {
    if (bp->cap) {
      NCX_HIP_CHECK(hipMemcpyAsync(ns.as<float>() + d * ncap, bp->stored.as<float>() + d * bp->cap, bp->cap * 4,
                                   hipMemcpyDeviceToDevice, ctx->stream));
      NCX_HIP_CHECK(hipMemcpyAsync(np.as<float>() + d * ncap, bp->pending.as<float>() + d * bp->cap, bp->cap * 4,
                                   hipMemcpyDeviceToDevice, ctx->stream));
    }
  }
  if (bp->cap) {
    NCX_HIP_CHECK(hipMemcpyAsync(nf.p, bp->pending_flag.p, bp->cap * 4, hipMemcpyDeviceToDevice, ctx->stream));
    NCX_HIP_CHECK(hipMemcpyAsync(nt.p, bp->in_tree.p, bp->cap * 4, hipMemcpyDeviceToDevice, ctx->stream));
  }
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  bp->stored.release(); bp->pending.release(); bp->pending_flag.release(); bp->in_tree.release();
  bp->stored = ns; bp->pending = np; bp->pending_flag = nf; bp->in_tree = nt;
  bp->cap = ncap;
  return NCX_OK;
}

extern "C" int ncx_bp_create(ncx_ctx* ctx, float margin, ncx_bp** out) {
  NCX_REQUIRE(ctx && out, "ncx_bp_create: null argument");
  NCX_REQUIRE(margin >= 0.f, "ncx_bp_create: the loosening margin must be positive");  // aabb.rs:192-195
  ncx_bp* bp = new ncx_bp;
  bp->ctx = ctx;
  bp->margin = margin;
  *out = bp;
  return NCX_OK;
}

extern "C" int ncx_bp_destroy(ncx_bp* bp) {
  if (!bp) return NCX_OK;
  DevBuf* bufs[] = {&bp->stored, &bp->pending, &bp->pending_flag, &bp->in_tree, &bp->alive, &bp->boxes,
                    &bp->cnt, &bp->offs, &bp->total, &bp->pa, &bp->pb, &bp->prev_a, &bp->prev_b,
                    &bp->sort_t1, &bp->sort_t2, &bp->sort_t3, &bp->sort_hist, &bp->sort_scan, &bp->ev_flags,
                    &bp->ev_offs, &bp->staging_h, &bp->staging_b, &bp->qcnt, &bp->qscan, &bp->qtot,
                    &bp->groups};
  for (DevBuf* b : bufs) b->release();
  bp->bvh.release();
  delete bp;
  return NCX_OK;
}

static uint32_t slab_insert(ncx_bp* bp) {
  uint32_t key = bp->slab_next;
  if (key == bp->occupied.size()) {
    bp->occupied.push_back(1);
    bp->next_free.push_back(0);
    bp->slab_next = key + 1;
  } else {
    bp->slab_next = bp->next_free[key];
    bp->occupied[key] = 1;
  }
  bp->slab_len++;
  if (bp->groups_on) {  // a new proxy starts with CollisionGroups::new()
    if (bp->groups_host.size() < 3 * bp->occupied.size()) bp->groups_host.resize(3 * bp->occupied.size(), 0u);
    bp->groups_host[3 * key] = NCX_ALL_GROUPS;
    bp->groups_host[3 * key + 1] = NCX_ALL_GROUPS;
    bp->groups_host[3 * key + 2] = 0u;
    bp->groups_dirty = true;
  }
  return key;
}

static void slab_remove(ncx_bp* bp, uint32_t key) {
  bp->occupied[key] = 0;
  bp->next_free[key] = bp->slab_next;
  bp->slab_next = key;
  bp->slab_len--;
}

// Applies queued boxes: mode 1 = create_proxy (stored as given), mode 2 =
// deferred_set_bounding_volume (containment test against the stored box, then loosened).
__global__ void k_bp_stage(const uint32_t* __restrict__ handles, const float* __restrict__ b0,
                           const float* __restrict__ b1, const float* __restrict__ b2, const float* __restrict__ b3,
                           const float* __restrict__ b4, const float* __restrict__ b5, size_t n, uint32_t mode,
                           float margin, const float* __restrict__ stored, float* __restrict__ pending,
                           uint32_t* __restrict__ pflag, const uint32_t* __restrict__ in_tree, size_t cap) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t h = handles[i];
  if (h >= cap) return;  // not a proxy slot: ignored (the host entry point rejects it up front)
  float b[6] = {b0[i], b1[i], b2[i], b3[i], b4[i], b5[i]};
  if (mode == 1) {
#pragma unroll
    for (int d = 0; d < 6; ++d) pending[d * cap + h] = b[d];
    pflag[h] = 1u;
    return;
  }
  bool needs = true;
  if (in_tree[h]) {
    // AABB::contains (aabb.rs:161-163)
    const float* s = stored;
    needs = !(s[0 * cap + h] <= b[0] && s[1 * cap + h] <= b[1] && s[2 * cap + h] <= b[2] &&
              s[3 * cap + h] >= b[3] && s[4 * cap + h] >= b[4] && s[5 * cap + h] >= b[5]);
  }
  if (needs) {
    // AABB::loosened (aabb.rs:190-199): mins + (-m), maxs + m
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      pending[d * cap + h] = b[d] + (-margin);
      pending[(3 + d) * cap + h] = b[3 + d] + margin;
    }
    pflag[h] = 2u;
  }
}

__global__ void k_bp_commit(const uint32_t* __restrict__ alive, size_t n, float* __restrict__ stored,
                            const float* __restrict__ pending, uint32_t* __restrict__ pflag,
                            uint32_t* __restrict__ in_tree, size_t cap, float* __restrict__ boxes,
                            uint32_t* __restrict__ changed) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t h = alive[i];
  if (pflag[h]) {
#pragma unroll
    for (int d = 0; d < 6; ++d) stored[d * cap + h] = pending[d * cap + h];
    pflag[h] = 0u;
    *changed = 1u;
  }
  in_tree[h] = 1u;
#pragma unroll
  for (int d = 0; d < 6; ++d) boxes[d * n + i] = stored[d * cap + h];
}

__device__ __forceinline__ bool box_overlap(const float q[6], float mnx, float mny, float mnz, float mxx, float mxy,
                                            float mxz) {
  // AABB::intersects (aabb.rs:156-158): mins <= other.maxs && maxs >= other.mins
  return mnx <= q[3] && mny <= q[4] && mnz <= q[5] && mxx >= q[0] && mxy >= q[1] && mxz >= q[2];
}

// Self-query of leaf k (sorted position) against leaves at positions > k.
// WRITE = false: count; WRITE = true: emit at offs[k].
template <bool WRITE>
__global__ __launch_bounds__(128) void k_bp_query(const LNode* __restrict__ nodes, const uint32_t* __restrict__ order,
                                                  const uint32_t* __restrict__ alive, const float* __restrict__ boxes,
                                                  uint32_t n, uint32_t* __restrict__ cnt,
                                                  const uint32_t* __restrict__ offs, uint32_t* __restrict__ pa,
                                                  uint32_t* __restrict__ pb, const uint32_t* __restrict__ groups) {
  __shared__ uint32_t stack[32 * 128];
  const int tid = threadIdx.x;
  uint32_t k = blockIdx.x * 128 + tid;
  if (k >= n) return;
  uint32_t prim = order[k];
  float q[6];
#pragma unroll
  for (int d = 0; d < 6; ++d) q[d] = boxes[(size_t)d * n + prim];
  uint32_t hk = alive[prim];
  uint32_t c = 0, w = WRITE ? offs[k] : 0u;
  auto emit = [&](uint32_t leafpos) {
    uint32_t h2 = alive[order[leafpos]];
    if (!groups_allow(groups, hk, h2)) return;  // is_interference_allowed (broad_phase.rs:31)
    if (WRITE) {
      pa[w] = hk < h2 ? hk : h2;
      pb[w] = hk < h2 ? h2 : hk;
      w++;
    } else {
      c++;
    }
  };
  int sp = 0;
  uint32_t node = 0;
  for (;;) {
    const float4* nq = (const float4*)(nodes + node);
    float4 q0 = nq[0], q1 = nq[1], q2 = nq[2], q3 = nq[3];
    uint32_t cl = __float_as_uint(q3.x), cr = __float_as_uint(q3.y);
    uint32_t ll = __float_as_uint(q3.z), lr = __float_as_uint(q3.w);
    bool hl = ll > k && box_overlap(q, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y);
    bool hr = lr > k && box_overlap(q, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w);
    if (hl && (cl & NCX_LEAF_BIT)) { emit(cl & ~NCX_LEAF_BIT); hl = false; }
    if (hr && (cr & NCX_LEAF_BIT)) { emit(cr & ~NCX_LEAF_BIT); hr = false; }
    if (hl && hr) {
      if (sp < 32) stack[sp * 128 + tid] = cr;
      sp++;  // depth beyond 32 is handled by the caller's depth check
      node = cl;
    } else if (hl) node = cl;
    else if (hr) node = cr;
    else {
      if (sp == 0) break;
      sp--;
      node = stack[sp * 128 + tid];
    }
  }
  if (!WRITE) cnt[k] = c;
}

// Generic-depth variant: a global stack per thread (used when the tree is deeper than the LDS stack).
template <bool WRITE>
__global__ __launch_bounds__(128) void k_bp_query_deep(const LNode* __restrict__ nodes,
                                                       const uint32_t* __restrict__ order,
                                                       const uint32_t* __restrict__ alive,
                                                       const float* __restrict__ boxes, uint32_t n,
                                                       uint32_t* __restrict__ cnt, const uint32_t* __restrict__ offs,
                                                       uint32_t* __restrict__ pa, uint32_t* __restrict__ pb,
                                                       uint32_t* __restrict__ gstack_all, uint32_t depth,
                                                       const uint32_t* __restrict__ groups) {
  uint32_t k = blockIdx.x * 128 + threadIdx.x;
  if (k >= n) return;
  uint32_t* stack = gstack_all + (size_t)k * depth;
  uint32_t prim = order[k];
  float q[6];
#pragma unroll
  for (int d = 0; d < 6; ++d) q[d] = boxes[(size_t)d * n + prim];
  uint32_t hk = alive[prim];
  uint32_t c = 0, w = WRITE ? offs[k] : 0u;
  int sp = 0;
  uint32_t node = 0;
  for (;;) {
    const LNode nd = nodes[node];
    bool hl = nd.last_l > k && box_overlap(q, nd.lminx, nd.lminy, nd.lminz, nd.lmaxx, nd.lmaxy, nd.lmaxz);
    bool hr = nd.last_r > k && box_overlap(q, nd.rminx, nd.rminy, nd.rminz, nd.rmaxx, nd.rmaxy, nd.rmaxz);
    uint32_t cl = nd.child_l, cr = nd.child_r;
    for (int side = 0; side < 2; ++side) {
      uint32_t ch = side ? cr : cl;
      bool h = side ? hr : hl;
      if (h && (ch & NCX_LEAF_BIT)) {
        uint32_t h2 = alive[order[ch & ~NCX_LEAF_BIT]];
        if (groups_allow(groups, hk, h2)) {
          if (WRITE) {
            pa[w] = hk < h2 ? hk : h2;
            pb[w] = hk < h2 ? h2 : hk;
            w++;
          } else c++;
        }
        if (side) hr = false; else hl = false;
      }
    }
    if (hl && hr) { stack[sp++] = cr; node = cl; }
    else if (hl) node = cl;
    else if (hr) node = cr;
    else {
      if (sp == 0) break;
      node = stack[--sp];
    }
  }
  if (!WRITE) cnt[k] = c;
}

__global__ void k_tree_depth(const uint32_t* __restrict__ parent, uint32_t n, uint32_t* __restrict__ maxdepth) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint32_t d = 1;
  uint32_t p = parent[(n - 1) + k] & NCX_PARENT_MASK;
  while (p != 0 && d <= n) {
    p = parent[p] & NCX_PARENT_MASK;
    d++;
  }
  atomicMax(maxdepth, d);
}

extern "C" int ncx_bp_create_proxies(ncx_bp* bp, const float* const aabb[6], size_t n, uint32_t* handles_out) {
  NCX_REQUIRE(bp && aabb && handles_out, "ncx_bp_create_proxies: null argument");
  if (n == 0) return NCX_OK;
  for (int d = 0; d < 6; ++d) NCX_REQUIRE(aabb[d], "ncx_bp_create_proxies: null box array");
  std::vector<uint32_t> hs(n);
  uint32_t maxh = 0;
  for (size_t i = 0; i < n; ++i) { hs[i] = slab_insert(bp); maxh = std::max(maxh, hs[i]); }
  int rc;
  if ((rc = bp_grow(bp, (size_t)maxh + 1))) return rc;
  if ((rc = bp->staging_h.reserve(n * 4)) || (rc = bp->staging_b.reserve(n * 24))) return rc;
  ncx_ctx* ctx = bp->ctx;
  NCX_HIP_CHECK(hipMemcpyAsync(bp->staging_h.p, hs.data(), n * 4, hipMemcpyHostToDevice, ctx->stream));
  float* sb = bp->staging_b.as<float>();
  for (int d = 0; d < 6; ++d)
    NCX_HIP_CHECK(hipMemcpyAsync(sb + d * n, aabb[d], n * 4, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_bp_stage, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, bp->staging_h.as<uint32_t>(), sb,
                     sb + n, sb + 2 * n, sb + 3 * n, sb + 4 * n, sb + 5 * n, n, 1u, bp->margin, bp->stored.as<float>(),
                     bp->pending.as<float>(), bp->pending_flag.as<uint32_t>(), bp->in_tree.as<uint32_t>(), bp->cap);
  NCX_HIP_CHECK(hipGetLastError());
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  for (size_t i = 0; i < n; ++i) handles_out[i] = hs[i];
  bp->topology_dirty = true;
  return NCX_OK;
}

static int check_handles(ncx_bp* bp, const uint32_t* hs, size_t n) {
  for (size_t i = 0; i < n; ++i)
    NCX_REQUIRE(hs[i] < bp->occupied.size() && bp->occupied[hs[i]],
                "Attempting to set the bounding volume of an object that does not exist.");
  return NCX_OK;
}

extern "C" int ncx_bp_set_aabbs_dev(ncx_bp* bp, const uint32_t* handles_dev, const float* const aabb[6], size_t n) {
  NCX_REQUIRE(bp && handles_dev && aabb, "ncx_bp_set_aabbs_dev: null argument");
  if (n == 0) return NCX_OK;
  for (int d = 0; d < 6; ++d) NCX_REQUIRE(aabb[d], "ncx_bp_set_aabbs_dev: null box array");
  ncx_ctx* ctx = bp->ctx;
  hipLaunchKernelGGL(k_bp_stage, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, handles_dev, aabb[0], aabb[1],
                     aabb[2], aabb[3], aabb[4], aabb[5], n, 2u, bp->margin, bp->stored.as<float>(),
                     bp->pending.as<float>(), bp->pending_flag.as<uint32_t>(), bp->in_tree.as<uint32_t>(), bp->cap);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

extern "C" int ncx_bp_set_aabbs(ncx_bp* bp, const uint32_t* handles, const float* const aabb[6], size_t n) {
  NCX_REQUIRE(bp && handles && aabb, "ncx_bp_set_aabbs: null argument");
  if (n == 0) return NCX_OK;
  int rc;
  if ((rc = check_handles(bp, handles, n))) return rc;
  if ((rc = bp->staging_h.reserve(n * 4)) || (rc = bp->staging_b.reserve(n * 24))) return rc;
  ncx_ctx* ctx = bp->ctx;
  NCX_HIP_CHECK(hipMemcpyAsync(bp->staging_h.p, handles, n * 4, hipMemcpyHostToDevice, ctx->stream));
  float* sb = bp->staging_b.as<float>();
  const float* dev[6];
  for (int d = 0; d < 6; ++d) {
    NCX_HIP_CHECK(hipMemcpyAsync(sb + d * n, aabb[d], n * 4, hipMemcpyHostToDevice, ctx->stream));
    dev[d] = sb + d * n;
  }
  rc = ncx_bp_set_aabbs_dev(bp, bp->staging_h.as<uint32_t>(), dev, n);
  if (rc) return rc;
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return NCX_OK;
}

static int bp_sort_current(ncx_bp* bp, bool prev);
static int download_keys(ncx_bp* bp, const DevBuf& a, const DevBuf& b, size_t n, uint64_t* out);

// BroadPhase::remove (dbvt_broad_phase.rs:282-323): every pair of the current set that contains a
// removed proxy is dropped and handed to the caller (the reference's handler), sorted.
extern "C" int ncx_bp_remove_with_pairs(ncx_bp* bp, const uint32_t* handles, size_t n, uint64_t* removed_keys,
                                        size_t cap, size_t* n_removed) {
  NCX_REQUIRE(bp && (handles || n == 0), "ncx_bp_remove: null argument");
  for (size_t i = 0; i < n; ++i)
    NCX_REQUIRE(handles[i] < bp->occupied.size() && bp->occupied[handles[i]],
                "Attempting to remove an object that does not exist.");
  std::vector<uint64_t> dropped;
  if (bp->npairs) {
    int rc;
    if ((rc = bp_sort_current(bp, false))) return rc;
    std::vector<uint64_t> cur(bp->npairs);
    if ((rc = download_keys(bp, bp->pa, bp->pb, bp->npairs, cur.data()))) return rc;
    std::vector<uint8_t> gone(bp->occupied.size(), 0);
    for (size_t i = 0; i < n; ++i) gone[handles[i]] = 1;
    std::vector<uint64_t> already = bp->removed_pending;
    std::sort(already.begin(), already.end());
    for (uint64_t k : cur) {
      const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
      if ((gone[a] || gone[b]) && !std::binary_search(already.begin(), already.end(), k)) dropped.push_back(k);
    }
  }
  if (n_removed) *n_removed = dropped.size();
  if (removed_keys && cap < dropped.size()) {
    ncx_set_error("ncx_bp_remove_with_pairs: capacity too small (nothing removed)");
    return NCX_E_CAPACITY;
  }
  if (removed_keys) std::copy(dropped.begin(), dropped.end(), removed_keys);
  bp->removed_pending.insert(bp->removed_pending.end(), dropped.begin(), dropped.end());
  ncx_ctx* ctx = bp->ctx;
  std::vector<uint32_t> zero(1, 0);
  for (size_t i = 0; i < n; ++i) {
    uint32_t h = handles[i];
    if (!bp->occupied[h]) continue;  // duplicate in the same call
    slab_remove(bp, h);
    NCX_HIP_CHECK(hipMemcpyAsync(bp->in_tree.as<uint32_t>() + h, zero.data(), 4, hipMemcpyHostToDevice, ctx->stream));
    NCX_HIP_CHECK(hipMemcpyAsync(bp->pending_flag.as<uint32_t>() + h, zero.data(), 4, hipMemcpyHostToDevice, ctx->stream));
  }
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  bp->topology_dirty = true;
  return NCX_OK;
}

extern "C" int ncx_bp_remove(ncx_bp* bp, const uint32_t* handles, size_t n) {
  return ncx_bp_remove_with_pairs(bp, handles, n, nullptr, 0, nullptr);
}

// Per-proxy CollisionGroups (src/pipeline/object/collision_groups.rs): the pair filter of
// BroadPhaseInterferenceHandler::is_interference_allowed (broad_phase.rs:31) as the world applies
// it (can_interact_with_groups, collision_groups.rs:353-359).  Host arrays.
extern "C" int ncx_bp_set_collision_groups(ncx_bp* bp, const uint32_t* handles, const uint32_t* membership,
                                           const uint32_t* whitelist, const uint32_t* blacklist, size_t n) {
  NCX_REQUIRE(bp && (n == 0 || (handles && membership && whitelist && blacklist)),
              "ncx_bp_set_collision_groups: null argument");
  for (size_t i = 0; i < n; ++i) {
    NCX_REQUIRE(handles[i] < bp->occupied.size() && bp->occupied[handles[i]],
                "ncx_bp_set_collision_groups: no such proxy");
    NCX_REQUIRE(membership[i] <= NCX_ALL_GROUPS && blacklist[i] <= NCX_ALL_GROUPS,
                "ncx_bp_set_collision_groups: there are at most 30 groups");
  }
  if (!bp->groups_on) {
    bp->groups_on = true;
    bp->groups_host.assign(3 * bp->occupied.size(), 0u);
    for (size_t h = 0; h < bp->occupied.size(); ++h) {
      bp->groups_host[3 * h] = NCX_ALL_GROUPS;
      bp->groups_host[3 * h + 1] = NCX_ALL_GROUPS;
    }
  }
  for (size_t i = 0; i < n; ++i) {
    bp->groups_host[3 * handles[i]] = membership[i];
    bp->groups_host[3 * handles[i] + 1] = whitelist[i];
    bp->groups_host[3 * handles[i] + 2] = blacklist[i];
  }
  bp->groups_dirty = true;
  return NCX_OK;
}

static int bp_query(ncx_bp* bp, uint32_t n) {
  ncx_ctx* ctx = bp->ctx;
  int rc;
  if ((rc = bp->cnt.reserve((size_t)n * 4)) || (rc = bp->offs.reserve((size_t)n * 4)) || (rc = bp->total.reserve(16)))
    return rc;
  const LNode* nodes = bp->bvh.nodes.as<LNode>();
  const uint32_t* order = bp->bvh.order.as<uint32_t>();
  // Tree depth decides between the LDS stack and a global stack.
  uint32_t depth = 0;
  NCX_HIP_CHECK(hipMemcpyAsync(bp->total.as<uint32_t>() + 1, &depth, 4, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_tree_depth, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, bp->bvh.parent.as<uint32_t>(),
                     n, bp->total.as<uint32_t>() + 1);
  NCX_HIP_CHECK(hipMemcpyAsync(&depth, bp->total.as<uint32_t>() + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  bool deep = depth >= 32;
  DevBuf& gst = bp->sort_t3;
  if (deep && (rc = gst.reserve((size_t)n * (depth + 1) * 4))) return rc;
  unsigned nb = ceil_div(n, 128);
  if (!deep)
    hipLaunchKernelGGL(k_bp_query<false>, dim3(nb), dim3(128), 0, ctx->stream, nodes, order, bp->alive.as<uint32_t>(),
                       bp->boxes.as<float>(), n, bp->cnt.as<uint32_t>(), (const uint32_t*)nullptr, (uint32_t*)nullptr,
                       (uint32_t*)nullptr, bp->groups_on ? bp->groups.as<uint32_t>() : (const uint32_t*)nullptr);
  else
    hipLaunchKernelGGL(k_bp_query_deep<false>, dim3(nb), dim3(128), 0, ctx->stream, nodes, order,
                       bp->alive.as<uint32_t>(), bp->boxes.as<float>(), n, bp->cnt.as<uint32_t>(),
                       (const uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr, gst.as<uint32_t>(), depth + 1,
                       bp->groups_on ? bp->groups.as<uint32_t>() : (const uint32_t*)nullptr);
  NCX_HIP_CHECK(hipGetLastError());
  if ((rc = scan_exclusive_u32(ctx, bp->cnt.as<uint32_t>(), bp->offs.as<uint32_t>(), n, bp->sort_scan,
                               bp->total.as<uint32_t>())))
    return rc;
  uint32_t total = 0;
  NCX_HIP_CHECK(hipMemcpyAsync(&total, bp->total.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if ((rc = bp->pa.reserve((size_t)total * 4 + 4)) || (rc = bp->pb.reserve((size_t)total * 4 + 4))) return rc;
  if (!deep)
    hipLaunchKernelGGL(k_bp_query<true>, dim3(nb), dim3(128), 0, ctx->stream, nodes, order, bp->alive.as<uint32_t>(),
                       bp->boxes.as<float>(), n, (uint32_t*)nullptr, bp->offs.as<uint32_t>(), bp->pa.as<uint32_t>(),
                       bp->pb.as<uint32_t>(), bp->groups_on ? bp->groups.as<uint32_t>() : (const uint32_t*)nullptr);
  else
    hipLaunchKernelGGL(k_bp_query_deep<true>, dim3(nb), dim3(128), 0, ctx->stream, nodes, order,
                       bp->alive.as<uint32_t>(), bp->boxes.as<float>(), n, (uint32_t*)nullptr,
                       bp->offs.as<uint32_t>(), bp->pa.as<uint32_t>(), bp->pb.as<uint32_t>(), gst.as<uint32_t>(),
                       depth + 1, bp->groups_on ? bp->groups.as<uint32_t>() : (const uint32_t*)nullptr);
  NCX_HIP_CHECK(hipGetLastError());
  bp->npairs = total;
  return NCX_OK;
}

extern "C" int ncx_bp_update(ncx_bp* bp) {
  NCX_REQUIRE(bp, "ncx_bp_update: null argument");
  ncx_ctx* ctx = bp->ctx;
  int rc;
  // previous pair list becomes prev
  std::swap(bp->pa, bp->prev_a);
  std::swap(bp->pb, bp->prev_b);
  bp->prev_npairs = bp->npairs;
  bp->prev_sorted = bp->sorted;
  bp->have_prev = true;
  bp->npairs = 0;
  bp->sorted = false;
  bp->removed_prev.swap(bp->removed_pending);
  bp->removed_pending.clear();
  if (bp->groups_on && bp->groups_dirty) {
    if (bp->groups_host.size() < 3 * bp->occupied.size()) bp->groups_host.resize(3 * bp->occupied.size(), 0u);
    if ((rc = bp->groups.reserve(bp->groups_host.size() * 4 + 4))) return rc;
    if (!bp->groups_host.empty())
      NCX_HIP_CHECK(hipMemcpyAsync(bp->groups.p, bp->groups_host.data(), bp->groups_host.size() * 4,
                                   hipMemcpyHostToDevice, ctx->stream));
    bp->groups_dirty = false;
  }
  if (bp->topology_dirty) {
    bp->alive_host.clear();
    for (uint32_t h = 0; h < bp->occupied.size(); ++h)
      if (bp->occupied[h]) bp->alive_host.push_back(h);
    if ((rc = bp->alive.reserve(bp->alive_host.size() * 4 + 4))) return rc;
    if (!bp->alive_host.empty())
      NCX_HIP_CHECK(hipMemcpyAsync(bp->alive.p, bp->alive_host.data(), bp->alive_host.size() * 4,
                                   hipMemcpyHostToDevice, ctx->stream));
  }
  size_t n = bp->alive_host.size();
  if (n == 0) {
    bp->bvh.n = 0;
    bp->topology_dirty = false;
    NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return NCX_OK;
  }
  if ((rc = bp->boxes.reserve(n * 24)) || (rc = bp->total.reserve(16))) return rc;
  if ((rc = fill_u32(ctx, bp->total.as<uint32_t>() + 2, 1, 0u))) return rc;
  hipLaunchKernelGGL(k_bp_commit, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, bp->alive.as<uint32_t>(), n,
                     bp->stored.as<float>(), bp->pending.as<float>(), bp->pending_flag.as<uint32_t>(),
                     bp->in_tree.as<uint32_t>(), bp->cap, bp->boxes.as<float>(), bp->total.as<uint32_t>() + 2);
  NCX_HIP_CHECK(hipGetLastError());
  const float* box[6];
  for (int d = 0; d < 6; ++d) box[d] = bp->boxes.as<float>() + d * n;
  if (bp->topology_dirty || bp->bvh.n != n) {
    if ((rc = lbvh_build(ctx, bp->bvh, box, n, 0.f))) return rc;
  } else {
    if ((rc = lbvh_refit(ctx, bp->bvh, box, 0.f))) return rc;
  }
  bp->topology_dirty = false;
  if (n == 1) {
    bp->npairs = 0;
    NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return NCX_OK;
  }
  if ((rc = bp_query(bp, (uint32_t)n))) return rc;
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return NCX_OK;
}

extern "C" int ncx_bp_num_pairs(ncx_bp* bp, size_t* n) {
  NCX_REQUIRE(bp && n, "ncx_bp_num_pairs: null argument");
  *n = bp->npairs;
  return NCX_OK;
}

static int bp_sort_current(ncx_bp* bp, bool prev) {
  size_t n = prev ? bp->prev_npairs : bp->npairs;
  bool& s = prev ? bp->prev_sorted : bp->sorted;
  if (s || n <= 1) { s = true; return NCX_OK; }
  int bits = 1;
  while (bits < 32 && ((size_t)1 << bits) < bp->occupied.size()) bits++;
  DevBuf& a = prev ? bp->prev_a : bp->pa;
  DevBuf& b = prev ? bp->prev_b : bp->pb;
  int rc = radix_sort_pairs64(bp->ctx, a.as<uint32_t>(), b.as<uint32_t>(), n, bits, bits, bp->sort_t1, bp->sort_t2,
                              bp->sort_t3, bp->sort_hist, bp->sort_scan);
  if (rc) return rc;
  s = true;
  return NCX_OK;
}

extern "C" int ncx_bp_pairs_dev(ncx_bp* bp, const uint32_t** a, const uint32_t** b, size_t* n) {
  NCX_REQUIRE(bp && a && b && n, "ncx_bp_pairs_dev: null argument");
  *a = bp->pa.as<uint32_t>();
  *b = bp->pb.as<uint32_t>();
  *n = bp->npairs;
  return NCX_OK;
}

extern "C" int ncx_bp_copy_pairs_dev(ncx_bp* bp, uint32_t* a_dev, uint32_t* b_dev, size_t cap, size_t* n) {
  NCX_REQUIRE(bp && n, "ncx_bp_copy_pairs_dev: null argument");
  *n = bp->npairs;
  if (cap < bp->npairs) {
    ncx_set_error("ncx_bp_copy_pairs_dev: capacity too small");
    return NCX_E_CAPACITY;
  }
  if (bp->npairs) {
    NCX_REQUIRE(a_dev && b_dev, "ncx_bp_copy_pairs_dev: null output");
    NCX_HIP_CHECK(hipMemcpyAsync(a_dev, bp->pa.p, bp->npairs * 4, hipMemcpyDeviceToDevice, bp->ctx->stream));
    NCX_HIP_CHECK(hipMemcpyAsync(b_dev, bp->pb.p, bp->npairs * 4, hipMemcpyDeviceToDevice, bp->ctx->stream));
  }
  return NCX_OK;
}

static int download_keys(ncx_bp* bp, const DevBuf& a, const DevBuf& b, size_t n, uint64_t* out) {
  std::vector<uint32_t> ha(n), hb(n);
  ncx_ctx* ctx = bp->ctx;
  if (n) {
    NCX_HIP_CHECK(hipMemcpyAsync(ha.data(), a.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    NCX_HIP_CHECK(hipMemcpyAsync(hb.data(), b.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  }
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  for (size_t i = 0; i < n; ++i) out[i] = ((uint64_t)ha[i] << 32) | hb[i];
  return NCX_OK;
}

extern "C" int ncx_bp_pairs(ncx_bp* bp, uint64_t* keys, size_t cap, size_t* n) {
  NCX_REQUIRE(bp && n, "ncx_bp_pairs: null argument");
  *n = bp->npairs;
  if (cap < bp->npairs) {
    ncx_set_error("ncx_bp_pairs: capacity too small");
    return NCX_E_CAPACITY;
  }
  int rc = bp_sort_current(bp, false);
  if (rc) return rc;
  return download_keys(bp, bp->pa, bp->pb, bp->npairs, keys);
}

extern "C" int ncx_bp_events(ncx_bp* bp, uint64_t* started, size_t cap_s, size_t* n_s, uint64_t* stopped,
                             size_t cap_t, size_t* n_t) {
  NCX_REQUIRE(bp && n_s && n_t, "ncx_bp_events: null argument");
  int rc;
  if ((rc = bp_sort_current(bp, false))) return rc;
  if ((rc = bp_sort_current(bp, true))) return rc;
  std::vector<uint64_t> cur(bp->npairs), prev(bp->prev_npairs);
  if ((rc = download_keys(bp, bp->pa, bp->pb, bp->npairs, cur.data()))) return rc;
  if ((rc = download_keys(bp, bp->prev_a, bp->prev_b, bp->prev_npairs, prev.data()))) return rc;
  if (!bp->removed_prev.empty()) {  // reported by remove(), not again as stopped
    std::vector<uint64_t> rm = bp->removed_prev, kept;
    std::sort(rm.begin(), rm.end());
    std::set_difference(prev.begin(), prev.end(), rm.begin(), rm.end(), std::back_inserter(kept));
    prev.swap(kept);
  }
  std::vector<uint64_t> s, t;
  std::set_difference(cur.begin(), cur.end(), prev.begin(), prev.end(), std::back_inserter(s));
  std::set_difference(prev.begin(), prev.end(), cur.begin(), cur.end(), std::back_inserter(t));
  *n_s = s.size();
  *n_t = t.size();
  if (cap_s < s.size() || cap_t < t.size()) {
    ncx_set_error("ncx_bp_events: capacity too small");
    return NCX_E_CAPACITY;
  }
  for (size_t i = 0; i < s.size(); ++i) started[i] = s[i];
  for (size_t i = 0; i < t.size(); ++i) stopped[i] = t[i];
  return NCX_OK;
}

extern "C" int ncx_bp_proxy_aabb(ncx_bp* bp, uint32_t handle, float out[6]) {
  NCX_REQUIRE(bp && out, "ncx_bp_proxy_aabb: null argument");
  NCX_REQUIRE(handle < bp->occupied.size() && bp->occupied[handle], "ncx_bp_proxy_aabb: no such proxy");
  ncx_ctx* ctx = bp->ctx;
  uint32_t in_tree = 0;
  NCX_HIP_CHECK(hipMemcpyAsync(&in_tree, bp->in_tree.as<uint32_t>() + handle, 4, hipMemcpyDeviceToHost, ctx->stream));
  for (int d = 0; d < 6; ++d)
    NCX_HIP_CHECK(hipMemcpyAsync(&out[d], bp->stored.as<float>() + d * bp->cap + handle, 4, hipMemcpyDeviceToHost,
                                 ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  NCX_REQUIRE(in_tree, "ncx_bp_proxy_aabb: proxy not yet inserted by an update");
  return NCX_OK;
}

// ------------------------------------------------------------------ interference queries
// BroadPhase::interferences_with_bounding_volume / _with_point / _with_ray (dbvt_broad_phase.rs:
// 387-432) over the tree of the last update, excluding proxies removed since: every proxy whose
// stored box intersects the query box (aabb.rs:156-158), contains the point (aabb.rs:138-146), or
// is hit by the ray within max_toi (AABB::intersects_ray, solid; ray_aabb.rs:13-50).  Results per
// query are handles at offsets_dev[q] .. offsets_dev[q + 1] (offsets_dev: nq + 1 entries).
static int bp_interferences(ncx_bp* bp, int kind, const float* const* q, int ncomp, size_t nq, float max_toi,
                            uint32_t* offsets_dev, uint32_t* handles_dev, size_t cap, size_t* total) {
  NCX_REQUIRE(bp && q && offsets_dev && total, "ncx_bp_interferences: null argument");
  NCX_REQUIRE(nq < 0xFFFFFFFFull, "ncx_bp_interferences: too many queries");
  for (int d = 0; d < ncomp; ++d) NCX_REQUIRE(q[d], "ncx_bp_interferences: null query array");
  ncx_ctx* ctx = bp->ctx;
  const size_t n = bp->alive_host.size();
  if (n == 0 || bp->bvh.n != n || nq == 0) {
    *total = 0;
    int rc;
    if (nq && (rc = fill_u32(ctx, offsets_dev, nq + 1, 0u))) return rc;
    NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return NCX_OK;
  }
  NCX_REQUIRE(handles_dev || cap == 0, "ncx_bp_interferences: null output");
  BQArgs a = {};
  a.nodes = bp->bvh.nodes.as<LNode>();
  a.order = bp->bvh.order.as<uint32_t>();
  a.alive = bp->alive.as<uint32_t>();
  a.in_tree = bp->in_tree.as<uint32_t>();
  a.leaf_soa = bp->boxes.as<float>();
  a.nleaves = (uint32_t)n;
  for (int d = 0; d < ncomp; ++d) a.q[d] = q[d];
  a.nq = (uint32_t)nq;
  a.max_toi = max_toi;
  a.handles = handles_dev;
  return bq_collect_run(ctx, a, kind, offsets_dev, cap, total, bp->qcnt, bp->qscan, bp->qtot);
}

extern "C" int ncx_bp_interferences_with_aabbs(ncx_bp* bp, const float* const aabb_soa_dev[6], size_t nq,
                                               uint32_t* offsets_dev, uint32_t* handles_dev, size_t cap,
                                               size_t* total) {
  return bp_interferences(bp, BQ_BOX, aabb_soa_dev, 6, nq, 0.f, offsets_dev, handles_dev, cap, total);
}

extern "C" int ncx_bp_interferences_with_points(ncx_bp* bp, const float* const point_soa_dev[3], size_t nq,
                                                uint32_t* offsets_dev, uint32_t* handles_dev, size_t cap,
                                                size_t* total) {
  return bp_interferences(bp, BQ_POINT, point_soa_dev, 3, nq, 0.f, offsets_dev, handles_dev, cap, total);
}

extern "C" int ncx_bp_interferences_with_rays(ncx_bp* bp, const float* const ray_soa_dev[6], size_t nq,
                                              float max_toi, uint32_t* offsets_dev, uint32_t* handles_dev,
                                              size_t cap, size_t* total) {
  return bp_interferences(bp, BQ_RAY, ray_soa_dev, 6, nq, max_toi, offsets_dev, handles_dev, cap, total);
}
