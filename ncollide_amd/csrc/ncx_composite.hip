// ncx_composite.hip -- a TriMesh (composite shape) against a batch of convex shapes on the GPU.
//
// Contact: contact_composite_shape_shape (src/query/contact/contact_composite_shape_shape.rs:9-45).
// For each query shape, the mesh triangles whose box intersects the shape's box in the mesh frame,
// loosened by the prediction, are its candidates (BoundingVolumeInterferencesCollector over the
// mesh's tree); query::contact(mesh pose, triangle, shape pose, shape, prediction) runs on every
// candidate pair in the batched, tiered contact kernels (the triangles join the shape set as
// Triangle parts); the deepest contact of each query is kept.
// Distance: distance_composite_shape_shape (src/query/distance/distance_composite_shape_shape.rs:
// 9-35): a search of the mesh's tree ordered by the distance from the origin to the Minkowski sum
// of a node's box and the shape's box (CompositeShapeAgainstAnyDistanceVisitor, :56-101), skipping
// boxes no nearer than the best distance so far; query::distance (GJK) on each triangle reached.
//
// The tree is the mesh's LBVH (its node boxes grown slightly for the ray cast, which only keeps
// more nodes); at the leaves the triangle's own box (its vertices' min and max) decides, as the
// reference's BVT leaves do.  The reference keeps the first of equally deep contacts in its tree's
// visiting order; here the least triangle id wins such an exact tie.
#include "ncx_contact.h"

#define MC_BLOCK 128
#define MC_STACK (LBVH_MAX_DEPTH + 2)

void ncx_mesh::release_composite() {
  if (comp) {
    comp->shapes.release(); comp->qerr.release();
    contact_tiers_release(comp->tiers);
    delete comp;
    comp = nullptr;
  }
  DevBuf* bufs[] = {&cposes, &ccnt, &coffs, &cscan, &ctot, &ca, &cb, &chas, &cw1, &cw2, &cn, &cdepth, &citers};
  for (DevBuf* b : bufs) b->release();
}

// Triangle k of the mesh as the shape record n + k of the combined set.
__global__ void k_mc_tri_records(ncx_shape* __restrict__ out, size_t ntri) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ntri) return;
  ncx_shape s = {};
  s.type = SHAPE_TRIANGLE;
  s.hull_offset = (uint32_t)k;
  s.pad[0] = __float_as_uint(1.f);  // bounding radius for the pair order (a heuristic only)
  out[k] = s;
}

__global__ void k_mc_fill_pose(ncx_pose* __restrict__ out, size_t n, ncx_pose p) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = p;
}

struct MCArgs {
  const LNode* nodes;
  const float4* tris;
  uint32_t ntri;
  ncx_pose mesh_pose;
  const ncx_shape* shapes;
  Pool P;
  const ncx_pose* poses;
  const uint32_t* qshape;
  uint32_t nq, nshapes;
  float prediction;
  uint32_t* count;
  const uint32_t* offs;
  uint32_t* ca;
  uint32_t* cb;
  uint32_t* err;
};

__device__ __forceinline__ void tri_box(const float4* tris, uint32_t k, f3& mn, f3& mx) {
  const float4 a = tris[3 * (size_t)k], b = tris[3 * (size_t)k + 1], c = tris[3 * (size_t)k + 2];
  mn = f3{fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z)};
  mx = f3{fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z)};
}

__device__ __forceinline__ bool box_hit(f3 amn, f3 amx, f3 bmn, f3 bmx) {  // AABB::intersects (aabb.rs:156-158)
  return amn.x <= bmx.x && amn.y <= bmx.y && amn.z <= bmx.z && amx.x >= bmn.x && amx.y >= bmn.y && amx.z >= bmn.z;
}

// Candidate triangles of each query (one lane per query): count pass, then write pass.
template <bool WRITE>
__global__ __launch_bounds__(MC_BLOCK) void k_mc_candidates(MCArgs a) {
  const uint32_t q = blockIdx.x * MC_BLOCK + threadIdx.x;
  if (q >= a.nq) return;
  const uint32_t s = a.qshape[q];
  if (s >= a.nshapes) {
    atomicOr(a.err, 1u);
    if (!WRITE) a.count[q] = 0;
    return;
  }
  iso m1;
  m1.t = f3{a.mesh_pose.tx, a.mesh_pose.ty, a.mesh_pose.tz};
  m1.q = quat{a.mesh_pose.qi, a.mesh_pose.qj, a.mesh_pose.qk, a.mesh_pose.qw};
  const iso ls = iso_inv_mul(m1, load_pose(a.poses, s));  // ls_m2 = m1.inverse() * m2
  f3 bmn, bmx;
  shape_aabb_g<1>(load_shape(a.shapes, s), ls, a.P, 0, bmn, bmx);
  // AABB::loosened (aabb.rs:190-199)
  bmn = bmn + f3{-a.prediction, -a.prediction, -a.prediction};
  bmx = bmx + f3{a.prediction, a.prediction, a.prediction};
  uint32_t cnt = 0, slot = WRITE ? a.offs[q] : 0u;
  auto leaf = [&](uint32_t k) {
    f3 tmn, tmx;
    tri_box(a.tris, k, tmn, tmx);
    if (!box_hit(tmn, tmx, bmn, bmx)) return;
    if (WRITE) {
      a.ca[slot] = a.nshapes + k;  // the triangle's record in the combined set
      a.cb[slot] = s;
      ++slot;
    }
    ++cnt;
  };
  if (a.ntri == 1) {
    leaf(0);
  } else {
    uint32_t stk[MC_STACK];
    int sp = 0;
    uint32_t node = 0;
    for (;;) {
      const LNode nd = a.nodes[node];
      bool hl = box_hit(f3{nd.lminx, nd.lminy, nd.lminz}, f3{nd.lmaxx, nd.lmaxy, nd.lmaxz}, bmn, bmx);
      bool hr = box_hit(f3{nd.rminx, nd.rminy, nd.rminz}, f3{nd.rmaxx, nd.rmaxy, nd.rmaxz}, bmn, bmx);
      if (hl && (nd.child_l & NCX_LEAF_BIT)) { leaf(nd.child_l & ~NCX_LEAF_BIT); hl = false; }
      if (hr && (nd.child_r & NCX_LEAF_BIT)) { leaf(nd.child_r & ~NCX_LEAF_BIT); hr = false; }
      if (hl && hr) {
        if (sp == MC_STACK) { atomicOr(a.err, 2u); break; }
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
  if (!WRITE) a.count[q] = cnt;
}

// The deepest contact of each query over its candidates (contact_composite_shape_shape.rs:33-42:
// replace when deeper), ties to the least triangle id.
__global__ void k_mc_best(const uint32_t* __restrict__ offs, uint32_t nq, const uint32_t* __restrict__ ca,
                          uint32_t nshapes, const float4* __restrict__ tris, const uint8_t* __restrict__ chas,
                          const float* __restrict__ cw1, const float* __restrict__ cw2, const float* __restrict__ cn,
                          const float* __restrict__ cdepth, uint8_t* __restrict__ has, float* __restrict__ w1,
                          float* __restrict__ w2, float* __restrict__ nrm, float* __restrict__ depth,
                          uint32_t* __restrict__ face) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  uint32_t best = 0xFFFFFFFFu, best_face = 0xFFFFFFFFu;
  float best_d = 0.f;
  for (uint32_t j = offs[q]; j < offs[q + 1]; ++j) {
    if (chas[j] != 1) continue;
    const float d = cdepth[j];
    const uint32_t f = __float_as_uint(tris[3 * (size_t)(ca[j] - nshapes)].w);
    if (best == 0xFFFFFFFFu || d > best_d || (d == best_d && f < best_face)) {
      best = j;
      best_d = d;
      best_face = f;
    }
  }
  const bool h = best != 0xFFFFFFFFu;
  has[q] = h ? 1 : 0;
  for (int d = 0; d < 3; ++d) {
    w1[3 * (size_t)q + d] = h ? cw1[3 * (size_t)best + d] : 0.f;
    w2[3 * (size_t)q + d] = h ? cw2[3 * (size_t)best + d] : 0.f;
    nrm[3 * (size_t)q + d] = h ? cn[3 * (size_t)best + d] : 0.f;
  }
  depth[q] = h ? best_d : 0.f;
  if (face) face[q] = best_face;
}

// AABB::distance_to_point of the origin, solid (src/query/point/point_aabb.rs:138-150), for the
// Minkowski sum of box (mn, mx) and the shape's box: mins + shift + (-margin), maxs + shift + margin
// (distance_composite_shape_shape.rs:74-79).
__device__ __forceinline__ float msum_dist(f3 mn, f3 mx, f3 shift, f3 margin) {
  const f3 lo = (mn + shift) + (-margin), hi = (mx + shift) + margin;
  const f3 z = f3{0.f, 0.f, 0.f};
  const f3 mins_pt = lo - z, pt_maxs = z - hi;
  const f3 sh = f3{fmaxf(fmaxf(mins_pt.x, pt_maxs.x), 0.f), fmaxf(fmaxf(mins_pt.y, pt_maxs.y), 0.f),
                   fmaxf(fmaxf(mins_pt.z, pt_maxs.z), 0.f)};
  return norm(sh);
}

struct MDArgs {
  const LNode* nodes;
  const float4* tris;
  uint32_t ntri;
  ncx_pose mesh_pose;
  const ncx_shape* shapes;
  Pool P;
  const ncx_pose* poses;
  const uint32_t* qshape;
  uint32_t nq, nshapes;
  float* dist;
  uint32_t* err;
};

#define MD_G 2
#define MD_BLOCK 128

// Distance of each query (one group of MD_G lanes per query): nearest-first search of the tree,
// skipping boxes whose Minkowski-sum distance is not below the best distance so far.
template <int G>
__global__ __launch_bounds__(MD_BLOCK) void k_md(MDArgs a) {
  const int gl = threadIdx.x % G;
  const size_t group = ((size_t)blockIdx.x * MD_BLOCK + threadIdx.x) / G;
  const size_t ngroups = (size_t)gridDim.x * MD_BLOCK / G;
  iso m1;
  m1.t = f3{a.mesh_pose.tx, a.mesh_pose.ty, a.mesh_pose.tz};
  m1.q = quat{a.mesh_pose.qi, a.mesh_pose.qj, a.mesh_pose.qk, a.mesh_pose.qw};
  for (size_t q = group; q < a.nq; q += ngroups) {
    const uint32_t s = a.qshape[q];
    if (s >= a.nshapes) {
      if (gl == 0) {
        atomicOr(a.err, 1u);
        a.dist[q] = 0.f;
      }
      continue;
    }
    Ctx2 c;
    c.m1 = m1;
    c.m2 = load_pose(a.poses, s);
    c.g2 = load_shape(a.shapes, s);
    c.g1.type = SHAPE_TRIANGLE;
    c.g1.cnt = 0;
    c.g1.a = c.g1.b = c.g1.c = 0.f;
    const iso ls = iso_inv_mul(m1, c.m2);
    f3 bmn, bmx;
    shape_aabb_g<G>(c.g2, ls, a.P, gl, bmn, bmx);
    const f3 shift = -((bmn + bmx) * 0.5f);  // -ls_aabb2.center() (aabb.rs:95-97)
    const f3 margin = (bmx - bmn) * 0.5f;     // ls_aabb2.half_extents() (aabb.rs:101-104)
    float best = NCX_FLT_MAX;
    auto leaf = [&](uint32_t k) {
      f3 tmn, tmx;
      tri_box(a.tris, k, tmn, tmx);
      if (!(msum_dist(tmn, tmx, shift, margin) < best)) return;
      // query::distance(m1, triangle, m2, g2): distance_support_map_support_map (GJK from
      // m1.t - m2.t, distance_support_map_support_map.rs:25-53)
      c.g1.off = k;
      f3 dir;
      float nl;
      if (!try_normalize(c.m1.t - c.m2.t, NCX_FLT_EPS, &dir, &nl)) dir = f3{1.f, 0.f, 0.f};
      Simplex sx;
      sx_init(sx);
      sx_reset(sx, cso_from<G>(c, dir, a.P, gl));
      const GJKOut r = gjk<G>(c, NCX_FLT_MAX, sx, a.P, gl, 10000);
      const float d = r.status == GJK_S_CLOSEST ? norm(r.p1 - r.p2) : 0.f;
      if (d < best) best = d;
    };
    if (a.ntri == 1) {
      leaf(0);
    } else {
      uint32_t stk[MC_STACK];
      float stc[MC_STACK];
      int sp = 0;
      uint32_t node = 0;
      bool overflow = false;
      for (;;) {
        const LNode nd = a.nodes[node];
        const float cl = msum_dist(f3{nd.lminx, nd.lminy, nd.lminz}, f3{nd.lmaxx, nd.lmaxy, nd.lmaxz}, shift, margin);
        const float cr = msum_dist(f3{nd.rminx, nd.rminy, nd.rminz}, f3{nd.rmaxx, nd.rmaxy, nd.rmaxz}, shift, margin);
        // nearer child first; a leaf is handled as it is reached
        const bool lfirst = cl <= cr;
        const uint32_t c0 = lfirst ? nd.child_l : nd.child_r, c1 = lfirst ? nd.child_r : nd.child_l;
        const float k0 = lfirst ? cl : cr, k1 = lfirst ? cr : cl;
        bool h0 = k0 < best, h1 = k1 < best;
        if (h0 && (c0 & NCX_LEAF_BIT)) { leaf(c0 & ~NCX_LEAF_BIT); h0 = false; }
        if (h1 && (c1 & NCX_LEAF_BIT) && k1 < best) { leaf(c1 & ~NCX_LEAF_BIT); h1 = false; }
        h1 = h1 && k1 < best;
        if (h0 && h1) {
          if (sp == MC_STACK) { overflow = true; break; }
          stk[sp] = c1;
          stc[sp] = k1;
          ++sp;
          node = c0;
        } else if (h0) {
          node = c0;
        } else if (h1) {
          node = c1;
        } else {
          bool found = false;
          while (sp > 0) {
            --sp;
            if (stc[sp] < best) { node = stk[sp]; found = true; break; }
          }
          if (!found) break;
        }
      }
      if (overflow && gl == 0) atomicOr(a.err, 2u);
    }
    if (gl == 0) a.dist[q] = best;
  }
}

static int mc_check(ncx_mesh* m, ncx_shapes* s, const char* what) {
  NCX_REQUIRE(m && s, std::string(what) + ": null argument");
  NCX_REQUIRE(m->ctx->device == s->ctx->device, std::string(what) + ": mesh and shapes on different devices");
  NCX_REQUIRE(s->n + m->ntri < 0x7FFFFFFFull, std::string(what) + ": too many shapes and triangles");
  return NCX_OK;
}

static int mc_err(ncx_mesh* m, const char* what) {
  uint32_t err = 0;
  ncx_ctx* ctx = m->ctx;
  NCX_HIP_CHECK(hipMemcpyAsync(&err, m->ctot.as<uint32_t>() + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (err & 1u) {
    ncx_set_error(std::string(what) + ": a shape index is not a shape of the set");
    return NCX_E_INVALID;
  }
  if (err & 2u) {
    ncx_set_error(std::string(what) + ": traversal stack overflow");
    return NCX_E_CAPACITY;
  }
  return NCX_OK;
}

extern "C" int ncx_mesh_contact_batch(ncx_mesh* m, const ncx_pose* mesh_pose, ncx_shapes* s,
                                      const ncx_pose* poses_dev, const uint32_t* shape_dev, size_t n,
                                      float prediction, uint8_t* has_dev, float* world1_dev, float* world2_dev,
                                      float* normal_dev, float* depth_dev, uint32_t* face_dev) {
  int rc;
  if ((rc = mc_check(m, s, "ncx_mesh_contact_batch"))) return rc;
  NCX_REQUIRE(mesh_pose && poses_dev && shape_dev && has_dev && world1_dev && world2_dev && normal_dev && depth_dev,
              "ncx_mesh_contact_batch: null argument");
  NCX_REQUIRE(prediction >= 0.f, "ncx_mesh_contact_batch: negative prediction");
  NCX_REQUIRE(n < 0xFFFFFFFFull, "ncx_mesh_contact_batch: too many queries");
  if (n == 0) return NCX_OK;
  ncx_ctx* ctx = m->ctx;
  const size_t ns = s->n, nt = m->ntri, nc = ns + nt;
  // the combined shape set: the shapes, then the triangles as Triangle parts
  if (!m->comp) {
    m->comp = new ncx_shapes;
    m->comp->ctx = ctx;
  }
  ncx_shapes* C = m->comp;
  if ((rc = C->shapes.reserve(nc * sizeof(ncx_shape)))) return rc;
  if (ns) NCX_HIP_CHECK(hipMemcpyAsync(C->shapes.p, s->shapes.p, ns * sizeof(ncx_shape), hipMemcpyDeviceToDevice,
                                       ctx->stream));
  hipLaunchKernelGGL(k_mc_tri_records, dim3(ceil_div(nt, 256)), dim3(256), 0, ctx->stream,
                     C->shapes.as<ncx_shape>() + ns, nt);
  C->n = nc;
  C->npts = s->npts;
  C->hx = s->hx; C->hy = s->hy; C->hz = s->hz;  // borrowed (never released through C)
  C->tri = m->tris.as<float4>();
  if ((rc = m->cposes.reserve(nc * sizeof(ncx_pose)))) return rc;
  if (ns) NCX_HIP_CHECK(hipMemcpyAsync(m->cposes.p, poses_dev, ns * sizeof(ncx_pose), hipMemcpyDeviceToDevice,
                                       ctx->stream));
  hipLaunchKernelGGL(k_mc_fill_pose, dim3(ceil_div(nt, 256)), dim3(256), 0, ctx->stream,
                     m->cposes.as<ncx_pose>() + ns, nt, *mesh_pose);
  NCX_HIP_CHECK(hipGetLastError());
  // candidates
  if ((rc = m->ccnt.reserve(n * 4)) || (rc = m->coffs.reserve((n + 1) * 4)) || (rc = m->ctot.reserve(16))) return rc;
  if ((rc = fill_u32(ctx, m->ctot.as<uint32_t>(), 2, 0u))) return rc;
  MCArgs a = {};
  a.nodes = m->bvh.nodes.as<LNode>();
  a.tris = m->tris.as<float4>();
  a.ntri = (uint32_t)nt;
  a.mesh_pose = *mesh_pose;
  a.shapes = s->shapes.as<ncx_shape>();
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>(), m->tris.as<float4>()};
  a.poses = poses_dev;
  a.qshape = shape_dev;
  a.nq = (uint32_t)n;
  a.nshapes = (uint32_t)ns;
  a.prediction = prediction;
  a.count = m->ccnt.as<uint32_t>();
  a.err = m->ctot.as<uint32_t>() + 1;
  const unsigned nb = ceil_div(n, MC_BLOCK);
  hipLaunchKernelGGL((k_mc_candidates<false>), dim3(nb), dim3(MC_BLOCK), 0, ctx->stream, a);
  NCX_HIP_CHECK(hipGetLastError());
  if ((rc = scan_exclusive_u32(ctx, a.count, m->coffs.as<uint32_t>(), n, m->cscan, m->ctot.as<uint32_t>()))) return rc;
  NCX_HIP_CHECK(hipMemcpyAsync(m->coffs.as<uint32_t>() + n, m->ctot.p, 4, hipMemcpyDeviceToDevice, ctx->stream));
  if ((rc = mc_err(m, "ncx_mesh_contact_batch"))) return rc;
  uint32_t total = 0;
  NCX_HIP_CHECK(hipMemcpyAsync(&total, m->ctot.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  const size_t cap = std::max<size_t>(total, 1);
  if ((rc = m->ca.reserve(cap * 4)) || (rc = m->cb.reserve(cap * 4)) || (rc = m->chas.reserve(cap)) ||
      (rc = m->cw1.reserve(cap * 12)) || (rc = m->cw2.reserve(cap * 12)) || (rc = m->cn.reserve(cap * 12)) ||
      (rc = m->cdepth.reserve(cap * 4)) || (rc = m->citers.reserve(cap * 4)))
    return rc;
  a.offs = m->coffs.as<uint32_t>();
  a.ca = m->ca.as<uint32_t>();
  a.cb = m->cb.as<uint32_t>();
  hipLaunchKernelGGL((k_mc_candidates<true>), dim3(nb), dim3(MC_BLOCK), 0, ctx->stream, a);
  NCX_HIP_CHECK(hipGetLastError());
  if (total) {
    // query::contact(mesh pose, triangle, shape pose, shape, prediction) on every candidate
    if ((rc = contact_tiers_reserve(C->tiers, total))) return rc;
    if ((rc = contacts_launch_device_count(C, C->tiers, ctx, m->cposes.as<ncx_pose>(), m->ca.as<uint32_t>(),
                                           m->cb.as<uint32_t>(), m->ctot.as<uint32_t>(), total, prediction, nullptr,
                                           10000, m->chas.as<uint8_t>(), m->cw1.as<float>(), m->cw2.as<float>(),
                                           m->cn.as<float>(), m->cdepth.as<float>(), m->citers.as<int32_t>(), true)))
      return rc;
    NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    size_t large = 0;
    if ((rc = contacts_resolve_overflow(C->tiers, ctx, &large))) return rc;
  }
  hipLaunchKernelGGL(k_mc_best, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, m->coffs.as<uint32_t>(),
                     (uint32_t)n, m->ca.as<uint32_t>(), (uint32_t)ns, m->tris.as<float4>(), m->chas.as<uint8_t>(),
                     m->cw1.as<float>(), m->cw2.as<float>(), m->cn.as<float>(), m->cdepth.as<float>(), has_dev,
                     world1_dev, world2_dev, normal_dev, depth_dev, face_dev);
  NCX_HIP_CHECK(hipGetLastError());
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return NCX_OK;
}

extern "C" int ncx_mesh_distance_batch(ncx_mesh* m, const ncx_pose* mesh_pose, ncx_shapes* s,
                                       const ncx_pose* poses_dev, const uint32_t* shape_dev, size_t n,
                                       float* dist_dev) {
  int rc;
  if ((rc = mc_check(m, s, "ncx_mesh_distance_batch"))) return rc;
  NCX_REQUIRE(mesh_pose && poses_dev && shape_dev && dist_dev, "ncx_mesh_distance_batch: null argument");
  NCX_REQUIRE(n < 0xFFFFFFFFull, "ncx_mesh_distance_batch: too many queries");
  if (n == 0) return NCX_OK;
  ncx_ctx* ctx = m->ctx;
  if ((rc = m->ctot.reserve(16)) || (rc = fill_u32(ctx, m->ctot.as<uint32_t>(), 2, 0u))) return rc;
  MDArgs a = {};
  a.nodes = m->bvh.nodes.as<LNode>();
  a.tris = m->tris.as<float4>();
  a.ntri = (uint32_t)m->ntri;
  a.mesh_pose = *mesh_pose;
  a.shapes = s->shapes.as<ncx_shape>();
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>(), m->tris.as<float4>()};
  a.poses = poses_dev;
  a.qshape = shape_dev;
  a.nq = (uint32_t)n;
  a.nshapes = (uint32_t)s->n;
  a.dist = dist_dev;
  a.err = m->ctot.as<uint32_t>() + 1;
  const unsigned blocks =
      std::max(1u, std::min<unsigned>(ceil_div(n * MD_G, MD_BLOCK), (unsigned)ctx->num_cus * 16));
  hipLaunchKernelGGL(k_md<MD_G>, dim3(blocks), dim3(MD_BLOCK), 0, ctx->stream, a);
  NCX_HIP_CHECK(hipGetLastError());
  return mc_err(m, "ncx_mesh_distance_batch");
}
