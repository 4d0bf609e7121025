// ncx_toi.hip -- batched linear time of impact and batched shape ray casts on the GPU.
//
// query::time_of_impact (src/query/time_of_impact/time_of_impact.rs:66-158) for Ball, Cuboid,
// Capsule and ConvexHull under translational motion: the Ball-Ball closed form
// (time_of_impact_ball_ball.rs:7-61) and, for every other pair, GJK's ray cast on the Minkowski
// difference (gjk::directional_distance, src/query/algorithms/gjk.rs:214-365) as
// time_of_impact_support_map_support_map does (time_of_impact_support_map_support_map.rs:10-116).
// RayCast::toi_and_normal_with_ray of the same four shapes (ray_ball.rs, ray_cuboid.rs,
// ray_aabb.rs, ray_support_map.rs) over a batch of (shape, ray) pairs.
//
// One group of TOI_G lanes evaluates one pair (SR_G lanes one ray on a support-map shape); the
// lanes split ConvexHull support scans and run the same control flow.  Groups take queries in a grid-stride
// loop; the ray kernel sorts each wavefront's rays by shape kind first (k_shape_rays).  All
// arithmetic restates the reference in f32 with its operation order.
#include "ncx_gjk.h"

#define TOI_G 4
#define TOI_BS 256

struct TOIOut {
  int has;
  float toi;
  f3 w1, w2, n1, n2;
  int status;  // 0 OutOfIterations, 1 Converged, 2 Failed, 3 Penetrating (time_of_impact.rs:7-25)
};

template <int G>
__device__ __forceinline__ TOIOut toi_pair(const iso& m1, f3 v1, const ShapeD& g1, const iso& m2, f3 v2,
                                           const ShapeD& g2, float max_toi, float target, const Pool& P, int gl) {
  TOIOut o;
  o.has = 0;
  o.toi = 0.f;
  o.status = 0;
  o.w1 = o.w2 = o.n1 = o.n2 = f3{0.f, 0.f, 0.f};
  if (g1.type == NCX_BALL && g2.type == NCX_BALL) {
    // time_of_impact_ball_ball (time_of_impact_ball_ball.rs:7-61), then the frame change of
    // query::time_of_impact (time_of_impact.rs:75-98)
    const f3 vel = v1 - v2;
    const float r1 = g1.a, r2 = g2.a;
    const float rsum = r1 + r2;
    const float radius = rsum + target;
    const f3 center = m1.t + (-m2.t);
    const f3 ro = f3{0.f, 0.f, 0.f}, rd = -vel;
    bool inside;
    float t;
    if (!ray_toi_with_ball(center, radius, ro, rd, true, inside, t)) return o;
    if (t > max_toi) return o;
    const f3 dpt = (ro + rd * t) - center;
    f3 n1, n2, w1, w2;
    if (radius == 0.f) {
      n1 = f3{1.f, 0.f, 0.f}; n2 = n1; w1 = f3{0.f, 0.f, 0.f}; w2 = w1;
    } else {
      n1 = dpt / radius;
      n2 = -n1;
      w1 = n1 * r1;
      w2 = n2 * r2;
    }
    o.has = 1;
    o.toi = t;
    o.status = (inside && norm_sq(center) < rsum * rsum) ? 3 : 1;
    o.w1 = qrot_inv(m1.q, w1);
    o.w2 = qrot_inv(m2.q, w2);
    o.n1 = qrot_inv(m1.q, n1);
    o.n2 = qrot_inv(m2.q, n2);
    return o;
  }
  // time_of_impact_support_map_support_map (time_of_impact_support_map_support_map.rs:10-116)
  RCtx c;
  c.m1 = m1; c.m2 = m2; c.g1 = g1; c.g2 = g2;
  c.dil = target;
  c.dilated = target != 0.f;
  Simplex s;
  sx_init(s);
  float t;
  f3 normal;
  if (!minkowski_ray_cast<G>(c, f3{0.f, 0.f, 0.f}, v2 - v1, NCX_FLT_MAX, s, P, gl, t, normal)) return o;
  f3 w1 = f3{0.f, 0.f, 0.f}, w2 = w1;
  if (t != 0.f) gjk_result(s, s.dim == 3, w1, w2);  // gjk.rs:228-237
  if (t > max_toi) return o;
  o.has = 1;
  o.toi = t;
  if (target == 0.f) {
    o.n1 = inv_xform_vec(m1, normal);
    o.n2 = inv_xform_vec(m2, -normal);
    o.w1 = inv_xform_pt(m1, w1);
    o.w2 = inv_xform_pt(m2, w2);
    o.status = t == 0.f ? 3 : 1;
    return o;
  }
  int status = 1;
  if (t == 0.f) {
    // closest_points_support_map_support_map_with_params(.., target_distance, Some(normal))
    // (src/query/closest_points/closest_points_support_map_support_map.rs:40-67)
    Ctx2 c2;
    c2.m1 = m1; c2.m2 = m2; c2.g1 = g1; c2.g2 = g2;
    f3 d;
    float nl;
    if (!try_normalize(normal, NCX_FLT_EPS, &d, &nl)) d = f3{1.f, 0.f, 0.f};
    Simplex s2;
    sx_init(s2);
    sx_reset(s2, cso_from<G>(c2, d, P, gl));
    GJKOut r = gjk<G>(c2, target, s2, P, gl, 10000);
    if (r.status == GJK_S_CLOSEST) {
      const f3 nn = normalize(r.p2 - r.p1);
      o.n1 = inv_xform_vec(m1, nn);
      o.n2 = inv_xform_vec(m2, -nn);
      o.w1 = inv_xform_pt(m1, w1);
      o.w2 = inv_xform_pt(m2, w2);
      o.status = 1;
      return o;
    } else if (r.status == GJK_S_INTERSECTION) {
      status = 3;
    }
  }
  const f3 n1 = inv_xform_vec(m1, normal);
  o.n1 = n1;
  o.n2 = inv_xform_vec(m2, -normal);
  o.w1 = inv_xform_pt(m1, w1) - n1 * target;
  o.w2 = inv_xform_pt(m2, w2);
  o.status = status;
  return o;
}

struct TOIArgs {
  const ncx_shape* shapes;
  Pool P;
  const ncx_pose* poses;
  const float* vel;
  const uint32_t* ia;
  const uint32_t* ib;
  size_t n;
  uint32_t nshapes;
  float max_toi, target;
  uint8_t* has;
  float* toi;
  float* w1;
  float* w2;
  float* n1;
  float* n2;
  uint8_t* status;
  uint32_t* err;
};

__device__ __forceinline__ void store3(float* p, size_t k, f3 v) {
  p[3 * k] = v.x;
  p[3 * k + 1] = v.y;
  p[3 * k + 2] = v.z;
}

template <int G>
__global__ __launch_bounds__(TOI_BS) void k_toi(TOIArgs a) {
  const int gl = threadIdx.x % G;
  const size_t group = ((size_t)blockIdx.x * TOI_BS + threadIdx.x) / G;
  const size_t ngroups = (size_t)gridDim.x * TOI_BS / G;
  for (size_t k = group; k < a.n; k += ngroups) {
    const uint32_t i = a.ia[k], j = a.ib[k];
    TOIOut o;
    if (i >= a.nshapes || j >= a.nshapes) {
      if (gl == 0) atomicOr(a.err, 1u);
      o.has = 0; o.toi = 0.f; o.status = 0;
      o.w1 = o.w2 = o.n1 = o.n2 = f3{0.f, 0.f, 0.f};
    } else {
      const iso m1 = load_pose(a.poses, i), m2 = load_pose(a.poses, j);
      const f3 v1 = f3{a.vel[3 * i], a.vel[3 * i + 1], a.vel[3 * i + 2]};
      const f3 v2 = f3{a.vel[3 * j], a.vel[3 * j + 1], a.vel[3 * j + 2]};
      o = toi_pair<G>(m1, v1, load_shape(a.shapes, i), m2, v2, load_shape(a.shapes, j), a.max_toi, a.target, a.P,
                      gl);
    }
    if (gl == 0) {
      a.has[k] = (uint8_t)o.has;
      a.toi[k] = o.toi;
      store3(a.w1, k, o.w1);
      store3(a.w2, k, o.w2);
      store3(a.n1, k, o.n1);
      store3(a.n2, k, o.n2);
      a.status[k] = (uint8_t)o.status;
    }
  }
}

struct SRArgs {
  const ncx_shape* shapes;
  Pool P;
  const ncx_pose* poses;
  const uint32_t* idx;
  const float *ox, *oy, *oz, *dx, *dy, *dz;
  size_t n;
  uint32_t nshapes;
  float max_toi;
  int solid;
  float* toi;
  float* nrm;
  uint32_t* feat;
  uint32_t* err;
};

// Shape ray casts, compacted per wavefront.  A wavefront takes 64 consecutive rays at a time:
// each ray on a Ball or a Cuboid (closed forms) is cast by its own lane at once; the rays on
// support-map shapes are listed in LDS, ConvexHulls first and Capsules after, and are cast
// G lanes per ray, 64 / G at a time, so that each pass holds rays of one kind save at most one
// pass.  The outputs of every ray are what shape_ray_cast<G> gives it.
#define SR_WAVES (TOI_BS / 64)
#define SR_G 2  // lanes per support-map ray (1.67 ms for 1,048,576 scene rays; 4 lanes: 1.91 ms, 8: 2.55 ms)
template <int G>
__global__ __launch_bounds__(TOI_BS) void k_shape_rays(SRArgs a) {
  static_assert(TOI_BS % 64 == 0 && 64 % G == 0, "whole wavefronts of whole groups");
  __shared__ uint8_t list[SR_WAVES][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int gl = lane % G, gi = lane / G;
  const size_t nw = (size_t)gridDim.x * SR_WAVES;
  for (size_t base = ((size_t)blockIdx.x * SR_WAVES + wv) * 64; base < a.n; base += nw * 64) {
    const size_t k = base + (size_t)lane;
    int cls = 0;  // 0: done by this lane; 1: ConvexHull; 2: Capsule
    if (k < a.n) {
      const uint32_t i = a.idx[k];
      float t = -1.f;
      f3 nn = f3{0.f, 0.f, 0.f};
      uint32_t f = 0xFFFFFFFFu;
      if (i >= a.nshapes) {
        atomicOr(a.err, 1u);
      } else {
        const ShapeD g = load_shape(a.shapes, i);
        if (g.type == NCX_CONVEX_HULL) cls = 1;
        else if (g.type == NCX_CAPSULE) cls = 2;
        else {
          float tt;
          f3 n;
          uint32_t ff;
          if (shape_ray_cast_closed(g, load_pose(a.poses, i), f3{a.ox[k], a.oy[k], a.oz[k]},
                                    f3{a.dx[k], a.dy[k], a.dz[k]}, a.max_toi, a.solid != 0, tt, n, ff)) {
            t = tt;
            nn = n;
            f = ff;
          }
        }
      }
      if (cls == 0) {
        a.toi[k] = t;
        store3(a.nrm, k, nn);
        a.feat[k] = f;
      }
    }
    const uint64_t mh = __ballot(cls == 1), mc = __ballot(cls == 2);
    const uint32_t nh = (uint32_t)__popcll(mh), m = nh + (uint32_t)__popcll(mc);
    if (cls) {
      const uint64_t mm = cls == 1 ? mh : mc;
      const uint32_t rank =
          __builtin_amdgcn_mbcnt_hi((uint32_t)(mm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mm, 0u));
      list[wv][(cls == 1 ? 0u : nh) + rank] = (uint8_t)lane;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (uint32_t e0 = 0; e0 < m; e0 += 64 / G) {
      const uint32_t e = e0 + (uint32_t)gi;
      if (e < m) {
        const size_t kk = base + list[wv][e];
        const uint32_t i = a.idx[kk];  // a shape of the set: listed above
        float t = -1.f, tt;
        f3 nn = f3{0.f, 0.f, 0.f}, n;
        uint32_t f = 0xFFFFFFFFu, ff;
        if (shape_ray_cast<G>(load_shape(a.shapes, i), load_pose(a.poses, i), f3{a.ox[kk], a.oy[kk], a.oz[kk]},
                              f3{a.dx[kk], a.dy[kk], a.dz[kk]}, a.max_toi, a.solid != 0, a.P, gl, tt, n, ff)) {
          t = tt;
          nn = n;
          f = ff;
        }
        if (gl == 0) {
          a.toi[kk] = t;
          store3(a.nrm, kk, nn);
          a.feat[kk] = f;
        }
      }
    }
    // the list is read in full before the next window writes it
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
}

static int query_err_check(ncx_shapes* s, const char* what) {
  ncx_ctx* ctx = s->ctx;
  uint32_t err = 0;
  NCX_HIP_CHECK(hipMemcpyAsync(&err, s->qerr.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (err) {
    ncx_set_error(std::string(what) + ": a shape index is not a shape of the set");
    return NCX_E_INVALID;
  }
  return NCX_OK;
}

static unsigned query_blocks(ncx_ctx* ctx, size_t n, size_t per_block = TOI_BS / TOI_G) {
  return std::max(1u, std::min<unsigned>(ceil_div(n, per_block), (unsigned)ctx->num_cus * 16));
}

extern "C" int ncx_time_of_impact_batch(ncx_shapes* s, const ncx_pose* poses_dev, const float* vel_dev,
                                        const uint32_t* ia_dev, const uint32_t* ib_dev, size_t npairs, float max_toi,
                                        float target_distance, uint8_t* has_dev, float* toi_dev,
                                        float* witness1_dev, float* witness2_dev, float* normal1_dev,
                                        float* normal2_dev, uint8_t* status_dev) {
  NCX_REQUIRE(s && poses_dev && vel_dev && ia_dev && ib_dev && has_dev && toi_dev && witness1_dev && witness2_dev &&
                  normal1_dev && normal2_dev && status_dev,
              "ncx_time_of_impact_batch: null argument");
  NCX_REQUIRE(target_distance >= 0.f, "ncx_time_of_impact_batch: negative target distance");
  if (npairs == 0) return NCX_OK;
  ncx_ctx* ctx = s->ctx;
  int rc;
  if ((rc = s->qerr.reserve(4)) || (rc = fill_u32(ctx, s->qerr.as<uint32_t>(), 1, 0u))) return rc;
  TOIArgs a;
  a.shapes = s->shapes.as<ncx_shape>();
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>()};
  a.poses = poses_dev;
  a.vel = vel_dev;
  a.ia = ia_dev;
  a.ib = ib_dev;
  a.n = npairs;
  a.nshapes = (uint32_t)s->n;
  a.max_toi = max_toi;
  a.target = target_distance;
  a.has = has_dev;
  a.toi = toi_dev;
  a.w1 = witness1_dev;
  a.w2 = witness2_dev;
  a.n1 = normal1_dev;
  a.n2 = normal2_dev;
  a.status = status_dev;
  a.err = s->qerr.as<uint32_t>();
  hipLaunchKernelGGL(k_toi<TOI_G>, dim3(query_blocks(ctx, npairs)), dim3(TOI_BS), 0, ctx->stream, a);
  NCX_HIP_CHECK(hipGetLastError());
  return query_err_check(s, "ncx_time_of_impact_batch");
}

extern "C" int ncx_shapes_cast_rays(ncx_shapes* s, const ncx_pose* poses_dev, const uint32_t* shape_dev,
                                    const float* const ray_soa_dev[6], size_t n, float max_toi, int solid,
                                    float* toi_dev, float* normal_dev, uint32_t* feature_dev) {
  NCX_REQUIRE(s && poses_dev && shape_dev && ray_soa_dev && toi_dev && normal_dev && feature_dev,
              "ncx_shapes_cast_rays: null argument");
  for (int d = 0; d < 6; ++d) NCX_REQUIRE(ray_soa_dev[d], "ncx_shapes_cast_rays: null ray array");
  if (n == 0) return NCX_OK;
  ncx_ctx* ctx = s->ctx;
  int rc;
  if ((rc = s->qerr.reserve(4)) || (rc = fill_u32(ctx, s->qerr.as<uint32_t>(), 1, 0u))) return rc;
  SRArgs a;
  a.shapes = s->shapes.as<ncx_shape>();
  a.P = Pool{s->hx.as<float>(), s->hy.as<float>(), s->hz.as<float>()};
  a.poses = poses_dev;
  a.idx = shape_dev;
  a.ox = ray_soa_dev[0]; a.oy = ray_soa_dev[1]; a.oz = ray_soa_dev[2];
  a.dx = ray_soa_dev[3]; a.dy = ray_soa_dev[4]; a.dz = ray_soa_dev[5];
  a.n = n;
  a.nshapes = (uint32_t)s->n;
  a.max_toi = max_toi;
  a.solid = solid;
  a.toi = toi_dev;
  a.nrm = normal_dev;
  a.feat = feature_dev;
  a.err = s->qerr.as<uint32_t>();
  hipLaunchKernelGGL(k_shape_rays<SR_G>, dim3(query_blocks(ctx, n, TOI_BS)), dim3(TOI_BS), 0, ctx->stream, a);
  NCX_HIP_CHECK(hipGetLastError());
  return query_err_check(s, "ncx_shapes_cast_rays");
}
