// ncx_raycast.hip -- TriMesh ray casting on the GPU.
//
// The reference (src/query/ray/ray_trimesh.rs:23-50) runs a best-first search
// over a median-split BVT with a BinaryHeap per ray.  Here the triangles are
// put in a Morton-sorted LBVH (boxes grown outward by a few ulps so culling is
// conservative) and each ray runs a depth-first, nearest-child-first closest-hit
// traversal with a per-thread stack in LDS.  The triangle test is
// ray_intersection_with_triangle (src/query/ray/ray_triangle.rs:32-114) with the
// same operation order, so every per-triangle toi is bit-identical to the
// reference's; the hit kept is the (toi, feature) lexicographic minimum.
#include "ncx_internal.h"

struct ncx_mesh {
  ncx_ctx* ctx;
  size_t npoints = 0, ntri = 0;
  LBVH bvh;
  DevBuf tris;      // float4[3 * ntri] in leaf order: a (w = face id bits), b, c
  DevBuf boxes;     // 6 SoA arrays of triangle boxes (build input)
  DevBuf pts, idx;  // uploaded mesh
  DevBuf overflow;  // spill area for traversal stacks deeper than the LDS part
  DevBuf err;       // error flag (stack exhausted)
};

#define RC_BLOCK 128
#define RC_LDS_STACK 24
#define RC_GLOBAL_STACK 104
#define RC_GRID_BLOCKS 4096

__global__ __launch_bounds__(256) void k_tri_boxes(const float* __restrict__ pts, const uint32_t* __restrict__ idx,
                                                    size_t ntri, float* __restrict__ bx) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntri) return;
  const float* a = pts + 3 * idx[3 * i];
  const float* b = pts + 3 * idx[3 * i + 1];
  const float* c = pts + 3 * idx[3 * i + 2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {  // Triangle::local_bounding_volume (aabb_triangle.rs:27-41)
    bx[d * ntri + i] = fminf(fminf(a[d], b[d]), c[d]);
    bx[(3 + d) * ntri + i] = fmaxf(fmaxf(a[d], b[d]), c[d]);
  }
}

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
// parameter, or NaN on a miss (so every `t <= x` test on it is false).
__device__ __forceinline__ float slab(float mnx, float mny, float mnz, float mxx, float mxy, float mxz, f3 o, f3 inv,
                                      bool zx, bool zy, bool zz, float tmax0) {
  float tmin = 0.f, tmax = tmax0;
  if (zx) { if (o.x < mnx || o.x > mxx) return __builtin_nanf(""); }
  else { float a = (mnx - o.x) * inv.x, b = (mxx - o.x) * inv.x; tmin = fmaxf(tmin, fminf(a, b)); tmax = fminf(tmax, fmaxf(a, b)); }
  if (zy) { if (o.y < mny || o.y > mxy) return __builtin_nanf(""); }
  else { float a = (mny - o.y) * inv.y, b = (mxy - o.y) * inv.y; tmin = fmaxf(tmin, fminf(a, b)); tmax = fminf(tmax, fmaxf(a, b)); }
  if (zz) { if (o.z < mnz || o.z > mxz) return __builtin_nanf(""); }
  else { float a = (mnz - o.z) * inv.z, b = (mxz - o.z) * inv.z; tmin = fmaxf(tmin, fminf(a, b)); tmax = fminf(tmax, fmaxf(a, b)); }
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
  auto fail = [&](int code) { m->bvh.release(); delete m; return code; };
  if ((rc = m->pts.reserve(npoints * 12)) || (rc = m->idx.reserve(ntriangles * 12)) ||
      (rc = m->boxes.reserve(ntriangles * 24)) || (rc = m->tris.reserve(ntriangles * 48)) ||
      (rc = m->err.reserve(4)))
    return fail(rc);
  if (hipMemcpyAsync(m->pts.p, points_xyz, npoints * 12, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipMemcpyAsync(m->idx.p, indices, ntriangles * 12, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
    ncx_set_error("ncx_trimesh_create: upload failed");
    return fail(NCX_E_HIP);
  }
  float* bx = m->boxes.as<float>();
  hipLaunchKernelGGL(k_tri_boxes, dim3(ceil_div(ntriangles, 256)), dim3(256), 0, ctx->stream, m->pts.as<float>(),
                     m->idx.as<uint32_t>(), ntriangles, bx);
  const float* box[6];
  for (int d = 0; d < 6; ++d) box[d] = bx + d * ntriangles;
  // Grow node boxes by 2^-18 (1 + |x|): larger than the rounding of any slab or triangle toi.
  if ((rc = lbvh_build(ctx, m->bvh, box, ntriangles, 3.8146973e-06f))) return fail(rc);
  if ((rc = fill_u32(ctx, m->err.as<uint32_t>(), 1, 0u))) return fail(rc);
  hipLaunchKernelGGL(k_tri_reorder, dim3(ceil_div(ntriangles, 256)), dim3(256), 0, ctx->stream, m->pts.as<float>(),
                     m->idx.as<uint32_t>(), m->bvh.order.as<uint32_t>(), ntriangles, m->tris.as<float4>());
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
    ncx_set_error("ncx_trimesh_create: build failed");
    return fail(NCX_E_HIP);
  }
  *out = m;
  return NCX_OK;
}

extern "C" int ncx_trimesh_destroy(ncx_mesh* m) {
  if (!m) return NCX_OK;
  m->bvh.release();
  m->tris.release(); m->boxes.release(); m->pts.release(); m->idx.release(); m->overflow.release(); m->err.release();
  delete m;
  return NCX_OK;
}

static int cast_impl(ncx_mesh* m, const ncx_pose* pose, const float* const r[6], size_t n, float max_toi, float* toi,
                     float* nrm, uint32_t* feat, uint32_t* sn, uint32_t* st, bool stats, bool sync) {
  NCX_REQUIRE(m && pose && r, "ncx_trimesh_cast_rays: null argument");
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
  else
    hipLaunchKernelGGL(k_cast<false>, dim3(blocks), dim3(RC_BLOCK), 0, ctx->stream, m->bvh.nodes.as<LNode>(),
                       m->tris.as<float4>(), (uint32_t)m->ntri, (uint32_t)m->ntri, *pose, r[0], r[1], r[2], r[3], r[4],
                       r[5], n, max_toi, toi, nrm, feat, sn, st, m->overflow.as<uint32_t>(), m->err.as<uint32_t>());
  NCX_HIP_CHECK(hipGetLastError());
  if (!sync) return NCX_OK;
  return ncx_trimesh_check(m);
}

// The traversal-stack error flag is sticky: it is cleared at mesh creation and checked here.
extern "C" int ncx_trimesh_check(ncx_mesh* m) {
  NCX_REQUIRE(m, "ncx_trimesh_check: null argument");
  uint32_t err = 0;
  NCX_HIP_CHECK(hipMemcpyAsync(&err, m->err.p, 4, hipMemcpyDeviceToHost, m->ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
  if (err) {
    ncx_set_error("ncx_trimesh_cast_rays: traversal stack exhausted (tree deeper than 128 levels)");
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
