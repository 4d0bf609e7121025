// ncx_bvh_query.h -- batched LBVH queries with one query per lane: boxes, points and rays against
// the broad phase's stored boxes (BroadPhase::interferences_with_bounding_volume / _with_point /
// _with_ray, src/pipeline/broad_phase/dbvt_broad_phase.rs:387-432), and the world-level ray
// queries that also cast the ray on each proxy's shape (glue::interferences_with_ray and
// first_interference_with_ray, src/pipeline/glue/query.rs:13-66, 200-224).
//
// Each lane walks the tree depth-first with its own stack (the tree is at most LBVH_MAX_DEPTH
// levels deep, so the stack never holds more entries than that).  The collectors run twice: a
// count pass, then, after a scan, a write pass to each query's slots.  The first-hit search
// visits the nearer child first and skips boxes entered after the best hit so far; the hit kept
// is the least (toi, handle), so the result does not depend on the visiting order.
#pragma once
#include "ncx_gjk.h"

#define BQ_BOX 0
#define BQ_POINT 1
#define BQ_RAY 2
#define BQ_BLOCK 128
#define BQ_NONE 0xFFFFFFFFu

struct BQArgs {
  const LNode* nodes;
  const uint32_t* order;  // sorted leaf position -> primitive index
  const uint32_t* alive;  // primitive index -> handle (null: the identity)
  const uint32_t* in_tree;  // per handle: nonzero while the proxy is in the tree (null: all are)
  const float* leaf_soa;  // primitive boxes, 6 SoA arrays of stride nleaves (read when nleaves == 1)
  uint32_t nleaves;
  const float* q[6];      // queries: box (mins, maxs), point (x, y, z), ray (origin, dir)
  uint32_t nq;
  float max_toi;
  // shapes, for the world-level ray queries
  const ncx_shape* shapes;
  Pool P;
  const ncx_pose* poses;
  int cast_shapes;
  // outputs
  uint32_t* count;        // count pass: hits per query
  const uint32_t* offs;   // write pass: first slot of each query
  uint32_t* handles;
  float* toi;
  float* nrm;
  uint32_t* err;          // set when a traversal stack would overflow (a tree deeper than the bound)
  const uint32_t* perm;   // first-hit rays: group k takes ray perm[k] (null: ray k)
};

#define BQ_STACK (LBVH_MAX_DEPTH + 2)

// AABB::toi_with_ray with solid = true at the identity (src/query/ray/ray_aabb.rs:13-50).
__device__ __forceinline__ bool bq_ray_box(f3 o, f3 d, float max_toi, float mnx, float mny, float mnz, float mxx,
                                           float mxy, float mxz, float& toi) {
  float tmin = 0.f, tmax = max_toi;
  const float mn[3] = {mnx, mny, mnz}, mx[3] = {mxx, mxy, mxz};
  const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (dd[i] == 0.f) {
      if (oo[i] < mn[i] || oo[i] > mx[i]) return false;
    } else {
      const float denom = 1.f / dd[i];
      float tn = (mn[i] - oo[i]) * denom;
      float tf = (mx[i] - oo[i]) * denom;
      if (tn > tf) { float x = tn; tn = tf; tf = x; }
      tmin = fmaxf(tmin, tn);
      tmax = fminf(tmax, tf);
      if (tmin > tmax) return false;
    }
  }
  toi = tmin;
  return true;
}

// The interference test of query kind K against one box; for rays `cost` is the box's entry toi.
template <int K>
__device__ __forceinline__ bool bq_test(const float qv[6], float max_toi, float mnx, float mny, float mnz, float mxx,
                                        float mxy, float mxz, float& cost) {
  cost = 0.f;
  if (K == BQ_BOX)  // AABB::intersects (src/bounding_volume/aabb.rs:156-158)
    return mnx <= qv[3] && mny <= qv[4] && mnz <= qv[5] && mxx >= qv[0] && mxy >= qv[1] && mxz >= qv[2];
  if (K == BQ_POINT)  // AABB::contains_local_point (aabb.rs:138-146)
    return !(qv[0] < mnx || qv[0] > mxx || qv[1] < mny || qv[1] > mxy || qv[2] < mnz || qv[2] > mxz);
  return bq_ray_box(f3{qv[0], qv[1], qv[2]}, f3{qv[3], qv[4], qv[5]}, max_toi, mnx, mny, mnz, mxx, mxy, mxz, cost);
}

__device__ __forceinline__ uint32_t bq_handle(const BQArgs& a, uint32_t leaf_pos) {
  const uint32_t prim = a.order[leaf_pos];
  return a.alive ? a.alive[prim] : prim;
}

// Collects every leaf whose box passes the test (and, with cast_shapes, whose shape the ray hits,
// solid = true).  WRITE = false counts; WRITE = true writes handles (and toi/normal) from offs[q].
template <int K, bool WRITE>
__global__ __launch_bounds__(BQ_BLOCK) void k_bq_collect(BQArgs a) {
  const uint32_t qi = blockIdx.x * BQ_BLOCK + threadIdx.x;
  if (qi >= a.nq) return;
  const int nq = K == BQ_POINT ? 3 : 6;
  float qv[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int d = 0; d < nq; ++d) qv[d] = a.q[d][qi];
  uint32_t cnt = 0;
  uint32_t slot = WRITE ? a.offs[qi] : 0u;
  auto emit = [&](uint32_t leaf_pos) {
    const uint32_t h = bq_handle(a, leaf_pos);
    if (a.in_tree && !a.in_tree[h]) return;  // removed since the last update
    float t = 0.f;
    f3 n = f3{0.f, 0.f, 0.f};
    if (K == BQ_RAY && a.cast_shapes) {
      uint32_t feat;
      if (!shape_ray_cast<1>(load_shape(a.shapes, h), load_pose(a.poses, h), f3{qv[0], qv[1], qv[2]},
                             f3{qv[3], qv[4], qv[5]}, a.max_toi, true, a.P, 0, t, n, feat))
        return;
    }
    if (WRITE) {
      a.handles[slot] = h;
      if (a.toi) a.toi[slot] = t;
      if (a.nrm) {
        a.nrm[3 * (size_t)slot] = n.x;
        a.nrm[3 * (size_t)slot + 1] = n.y;
        a.nrm[3 * (size_t)slot + 2] = n.z;
      }
      ++slot;
    }
    ++cnt;
  };
  if (a.nleaves == 1) {
    const float* b = a.leaf_soa;
    float c;
    if (bq_test<K>(qv, a.max_toi, b[0], b[1], b[2], b[3], b[4], b[5], c)) emit(0);
  } else {
    uint32_t stk[BQ_STACK];
    int sp = 0;
    uint32_t node = 0;
    for (;;) {
      const LNode nd = a.nodes[node];
      float cl, cr;
      bool hl = bq_test<K>(qv, a.max_toi, nd.lminx, nd.lminy, nd.lminz, nd.lmaxx, nd.lmaxy, nd.lmaxz, cl);
      bool hr = bq_test<K>(qv, a.max_toi, nd.rminx, nd.rminy, nd.rminz, nd.rmaxx, nd.rmaxy, nd.rmaxz, cr);
      if (hl && (nd.child_l & NCX_LEAF_BIT)) { emit(nd.child_l & ~NCX_LEAF_BIT); hl = false; }
      if (hr && (nd.child_r & NCX_LEAF_BIT)) { emit(nd.child_r & ~NCX_LEAF_BIT); hr = false; }
      if (hl && hr) {
        if (sp == BQ_STACK) { atomicOr(a.err, 1u); break; }
        stk[sp++] = nd.child_r;
        node = nd.child_l;
      } else if (hl) {
        node = nd.child_l;
      } else if (hr) {
        node = nd.child_r;
      } else if (sp > 0) {
        node = stk[--sp];
      } else {
        break;
      }
    }
  }
  if (!WRITE) a.count[qi] = cnt;
}

// First hit of each ray among the proxies' shapes (solid = true, toi <= max_toi): the least
// (toi, handle) over the leaves whose box the ray enters no later than the best hit so far.
// G lanes per ray: they run the same search, and split the support scans of hull shapes at the
// leaves (shape_ray_cast<G>).
#define BQ_RAY_G 4
template <int G>
static __global__ __launch_bounds__(BQ_BLOCK) void k_bq_first_ray(BQArgs a) {
  const uint32_t gi = (blockIdx.x * BQ_BLOCK + threadIdx.x) / G;
  const int gl = threadIdx.x % G;
  if (gi >= a.nq) return;  // whole groups: BQ_BLOCK is a multiple of G
  const uint32_t qi = a.perm ? a.perm[gi] : gi;
  float qv[6];
  for (int d = 0; d < 6; ++d) qv[d] = a.q[d][qi];
  const f3 o = f3{qv[0], qv[1], qv[2]}, dir = f3{qv[3], qv[4], qv[5]};
  float best_t = a.max_toi;
  uint32_t best_h = BQ_NONE;
  f3 best_n = f3{0.f, 0.f, 0.f};
  auto leaf = [&](uint32_t leaf_pos) {
    const uint32_t h = bq_handle(a, leaf_pos);
    float t;
    f3 n;
    uint32_t feat;
    if (shape_ray_cast<G>(load_shape(a.shapes, h), load_pose(a.poses, h), o, dir, a.max_toi, true, a.P, gl, t, n,
                          feat)) {
      if (t < best_t || (t == best_t && h < best_h)) {
        best_t = t;
        best_h = h;
        best_n = n;
      }
    }
  };
  if (a.nleaves == 1) {
    const float* b = a.leaf_soa;
    float c;
    if (bq_ray_box(o, dir, a.max_toi, b[0], b[1], b[2], b[3], b[4], b[5], c)) leaf(0);
  } else {
    uint32_t stk[BQ_STACK];
    float stc[BQ_STACK];
    int sp = 0;
    uint32_t node = 0;
    for (;;) {
      const LNode nd = a.nodes[node];
      float cl, cr;
      bool hl = bq_ray_box(o, dir, a.max_toi, nd.lminx, nd.lminy, nd.lminz, nd.lmaxx, nd.lmaxy, nd.lmaxz, cl) &&
                cl <= best_t;
      bool hr = bq_ray_box(o, dir, a.max_toi, nd.rminx, nd.rminy, nd.rminz, nd.rmaxx, nd.rmaxy, nd.rmaxz, cr) &&
                cr <= best_t;
      if (hl && (nd.child_l & NCX_LEAF_BIT)) { leaf(nd.child_l & ~NCX_LEAF_BIT); hl = false; }
      if (hr && (nd.child_r & NCX_LEAF_BIT)) { leaf(nd.child_r & ~NCX_LEAF_BIT); hr = false; }
      hl = hl && cl <= best_t;
      hr = hr && cr <= best_t;
      if (hl && hr) {
        if (sp == BQ_STACK) { atomicOr(a.err, 1u); break; }
        const bool lfirst = cl <= cr;
        stk[sp] = lfirst ? nd.child_r : nd.child_l;
        stc[sp] = lfirst ? cr : cl;
        ++sp;
        node = lfirst ? nd.child_l : nd.child_r;
      } else if (hl) {
        node = nd.child_l;
      } else if (hr) {
        node = nd.child_r;
      } else {
        // pop entries until one can still hold a better hit
        bool found = false;
        while (sp > 0) {
          --sp;
          if (stc[sp] <= best_t) { node = stk[sp]; found = true; break; }
        }
        if (!found) break;
      }
    }
  }
  if (gl != 0) return;
  a.handles[qi] = best_h;
  a.toi[qi] = best_h == BQ_NONE ? -1.f : best_t;
  a.nrm[3 * (size_t)qi] = best_n.x;
  a.nrm[3 * (size_t)qi + 1] = best_n.y;
  a.nrm[3 * (size_t)qi + 2] = best_n.z;
}
static_assert(BQ_BLOCK % BQ_RAY_G == 0, "a block holds whole groups");

// Count pass, scan, write pass.  *total = hits over all queries; NCX_E_CAPACITY when above cap
// (nothing written).  offs (nq + 1 entries) receives each query's first slot and the total.
inline int bq_collect_run(ncx_ctx* ctx, BQArgs a, int kind, uint32_t* offs, size_t cap, size_t* total, DevBuf& cnt,
                          DevBuf& scan_tmp, DevBuf& tot) {
  if (a.nq == 0) {
    *total = 0;
    return NCX_OK;
  }
  int rc;
  if ((rc = cnt.reserve((size_t)a.nq * 4)) || (rc = tot.reserve(16))) return rc;
  if ((rc = fill_u32(ctx, tot.as<uint32_t>() + 1, 1, 0u))) return rc;
  a.count = cnt.as<uint32_t>();
  a.err = tot.as<uint32_t>() + 1;
  const unsigned nb = ceil_div(a.nq, BQ_BLOCK);
  if (kind == BQ_BOX) hipLaunchKernelGGL((k_bq_collect<BQ_BOX, false>), dim3(nb), dim3(BQ_BLOCK), 0, ctx->stream, a);
  else if (kind == BQ_POINT) hipLaunchKernelGGL((k_bq_collect<BQ_POINT, false>), dim3(nb), dim3(BQ_BLOCK), 0, ctx->stream, a);
  else hipLaunchKernelGGL((k_bq_collect<BQ_RAY, false>), dim3(nb), dim3(BQ_BLOCK), 0, ctx->stream, a);
  NCX_HIP_CHECK(hipGetLastError());
  if ((rc = scan_exclusive_u32(ctx, a.count, offs, a.nq, scan_tmp, tot.as<uint32_t>()))) return rc;
  uint32_t t[2] = {0, 0};
  NCX_HIP_CHECK(hipMemcpyAsync(t, tot.p, 8, hipMemcpyDeviceToHost, ctx->stream));
  NCX_HIP_CHECK(hipMemcpyAsync(offs + a.nq, tot.p, 4, hipMemcpyDeviceToDevice, ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  NCX_REQUIRE(t[1] == 0, "interference query: traversal stack overflow");
  *total = t[0];
  if (t[0] > cap) {
    ncx_set_error("interference query: " + std::to_string(t[0]) + " hits exceed the output capacity");
    return NCX_E_CAPACITY;
  }
  a.offs = offs;
  if (kind == BQ_BOX) hipLaunchKernelGGL((k_bq_collect<BQ_BOX, true>), dim3(nb), dim3(BQ_BLOCK), 0, ctx->stream, a);
  else if (kind == BQ_POINT) hipLaunchKernelGGL((k_bq_collect<BQ_POINT, true>), dim3(nb), dim3(BQ_BLOCK), 0, ctx->stream, a);
  else hipLaunchKernelGGL((k_bq_collect<BQ_RAY, true>), dim3(nb), dim3(BQ_BLOCK), 0, ctx->stream, a);
  NCX_HIP_CHECK(hipGetLastError());
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return NCX_OK;
}
