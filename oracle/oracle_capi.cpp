// oracle_capi.cpp -- extern "C" surface of the CPU oracle (ncx_oracle.hpp).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker; never by the product library.
//
// Layouts (all host memory, caller-owned):
//   shapes : types[n] (0 ball, 1 cuboid, 2 capsule, 3 hull), params[3n],
//            hull_off[n], hull_cnt[n] index into hull_pts[3 * total] (xyz AoS)
//   poses  : 7 per shape, (tx, ty, tz, qi, qj, qk, qw)
//   pairs  : ia[np], ib[np] shape indices
#include <memory>
#include <thread>
#include <vector>
#include <algorithm>
#include "ncx_oracle.hpp"

using namespace ncxo;

namespace {

template <class R> struct ShapeSet {
  std::vector<Shape<R>> shapes;
  std::vector<V3<R>> pts;
  void load(int64_t n, const uint32_t* types, const R* params, const uint32_t* hoff, const uint32_t* hcnt,
            const R* hpts, int64_t npts_total) {
    pts.resize((size_t)npts_total);
    for (int64_t i = 0; i < npts_total; ++i) pts[i] = V3<R>(hpts[3 * i], hpts[3 * i + 1], hpts[3 * i + 2]);
    shapes.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      Shape<R>& s = shapes[i];
      s.type = types[i];
      s.a = params[3 * i]; s.b = params[3 * i + 1]; s.c = params[3 * i + 2];
      s.pts = nullptr; s.npts = 0;
      if (s.type == HULL) { s.pts = pts.data() + hoff[i]; s.npts = hcnt[i]; }
    }
  }
};

template <class R> inline Iso<R> pose_at(const R* poses, int64_t i) {
  Iso<R> m;
  const R* p = poses + 7 * i;
  m.t = V3<R>(p[0], p[1], p[2]);
  m.q = Quat<R>{p[3], p[4], p[5], p[6]};
  return m;
}

template <class F> void parallel_for(int64_t n, int nthreads, F f) {
  if (nthreads <= 1 || n < 2) { for (int64_t i = 0; i < n; ++i) f(i); return; }
  std::vector<std::thread> th;
  int64_t chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=]() { for (int64_t i = b; i < e; ++i) f(i); });
  }
  for (auto& t : th) t.join();
}

template <class R>
int contact_batch(int64_t ns, const uint32_t* types, const R* params, const uint32_t* hoff, const uint32_t* hcnt,
                  const R* hpts, int64_t npts_total, const R* poses, const uint32_t* ia, const uint32_t* ib, int64_t np,
                  R prediction, uint8_t* has, R* w1, R* w2, R* nrm, R* depth, int32_t* gjk_it, int32_t* epa_it,
                  int nthreads) {
  ShapeSet<R> ss;
  ss.load(ns, types, params, hoff, hcnt, hpts, npts_total);
  parallel_for(np, nthreads, [&](int64_t k) {
    uint32_t a = ia[k], b = ib[k];
    Contact<R> c;
    int gi = 0, ei = 0;
    bool ok = contact(pose_at(poses, a), ss.shapes[a], pose_at(poses, b), ss.shapes[b], prediction, &c, &gi, &ei);
    has[k] = ok ? 1 : 0;
    if (gjk_it) gjk_it[k] = gi;
    if (epa_it) epa_it[k] = ei;
    if (ok) {
      for (int d = 0; d < 3; ++d) { w1[3 * k + d] = c.world1[d]; w2[3 * k + d] = c.world2[d]; nrm[3 * k + d] = c.normal[d]; }
      depth[k] = c.depth;
    } else {
      for (int d = 0; d < 3; ++d) { w1[3 * k + d] = 0; w2[3 * k + d] = 0; nrm[3 * k + d] = 0; }
      depth[k] = 0;
    }
  });
  return 0;
}

template <class R>
int contact_sm_batch(int64_t ns, const uint32_t* types, const R* params, const uint32_t* hoff, const uint32_t* hcnt,
                     const R* hpts, int64_t npts_total, const R* poses, const uint32_t* ia, const uint32_t* ib, int64_t np,
                     R prediction, const R* init_dirs, int32_t gjk_max_iter, uint8_t* has, R* w1, R* w2, R* nrm,
                     R* depth, int32_t* gjk_it, int32_t* epa_it, int32_t* status, int nthreads) {
  ShapeSet<R> ss;
  ss.load(ns, types, params, hoff, hcnt, hpts, npts_total);
  parallel_for(np, nthreads, [&](int64_t k) {
    uint32_t a = ia[k], b = ib[k];
    Contact<R> c;
    int gi = 0, ei = 0, st = 0;
    V3<R> d0;
    const V3<R>* dptr = nullptr;
    if (init_dirs) {
      d0 = V3<R>(init_dirs[3 * k], init_dirs[3 * k + 1], init_dirs[3 * k + 2]);
      if (d0.x != R(0) || d0.y != R(0) || d0.z != R(0)) dptr = &d0;  // all-zero entry = None
    }
    bool ok = contact_support_map_support_map(pose_at(poses, a), ss.shapes[a], pose_at(poses, b), ss.shapes[b],
                                              prediction, dptr, &c, &gi, &ei, dptr ? gjk_max_iter : 10000, &st);
    has[k] = ok ? 1 : 0;
    gjk_it[k] = gi; epa_it[k] = ei; status[k] = st;
    for (int d = 0; d < 3; ++d) {
      w1[3 * k + d] = ok ? c.world1[d] : R(0);
      w2[3 * k + d] = ok ? c.world2[d] : R(0);
      nrm[3 * k + d] = ok ? c.normal[d] : R(0);
    }
    depth[k] = ok ? c.depth : R(0);
  });
  return 0;
}

template <class R>
int distance_batch(int64_t ns, const uint32_t* types, const R* params, const uint32_t* hoff, const uint32_t* hcnt,
                   const R* hpts, int64_t npts_total, const R* poses, const uint32_t* ia, const uint32_t* ib, int64_t np,
                   const R* init_dirs, int32_t max_iter, R* dist, int32_t* gjk_it, int nthreads) {
  ShapeSet<R> ss;
  ss.load(ns, types, params, hoff, hcnt, hpts, npts_total);
  parallel_for(np, nthreads, [&](int64_t k) {
    uint32_t a = ia[k], b = ib[k];
    int gi = 0;
    V3<R> idir;
    const V3<R>* ip = nullptr;
    if (init_dirs) {  // an all-zero row means None
      idir = V3<R>(init_dirs[3 * k], init_dirs[3 * k + 1], init_dirs[3 * k + 2]);
      if (idir.x != 0 || idir.y != 0 || idir.z != 0) ip = &idir;
    }
    dist[k] = distance(pose_at(poses, a), ss.shapes[a], pose_at(poses, b), ss.shapes[b], ip, &gi, max_iter);
    if (gjk_it) gjk_it[k] = gi;
  });
  return 0;
}

template <class R>
int shape_aabbs(int64_t ns, const uint32_t* types, const R* params, const uint32_t* hoff, const uint32_t* hcnt,
                const R* hpts, int64_t npts_total, const R* poses, R* out6) {
  ShapeSet<R> ss;
  ss.load(ns, types, params, hoff, hcnt, hpts, npts_total);
  for (int64_t i = 0; i < ns; ++i) {
    AABB<R> bb = shape_aabb(ss.shapes[i], pose_at(poses, i));
    out6[6 * i + 0] = bb.mins.x; out6[6 * i + 1] = bb.mins.y; out6[6 * i + 2] = bb.mins.z;
    out6[6 * i + 3] = bb.maxs.x; out6[6 * i + 4] = bb.maxs.y; out6[6 * i + 5] = bb.maxs.z;
  }
  return 0;
}

template <class R> struct MeshHandle { TriMesh<R> mesh; };

template <class R> void* trimesh_create(const R* pts, int64_t npts, const uint32_t* idx, int64_t ntri) {
  auto* h = new MeshHandle<R>();
  h->mesh.points.resize((size_t)npts);
  for (int64_t i = 0; i < npts; ++i) h->mesh.points[i] = V3<R>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  h->mesh.idx.assign(idx, idx + 3 * ntri);
  h->mesh.build();
  return h;
}

template <class R>
int trimesh_cast(void* hp, const R* pose7, const R* orig, const R* dir, int64_t n, R max_toi, R* toi, R* nrm,
                 uint32_t* feature, uint32_t* visits, int nthreads) {
  auto* h = (MeshHandle<R>*)hp;
  Iso<R> m = pose_at(pose7, 0);
  parallel_for(n, nthreads, [&](int64_t k) {
    Ray<R> ray{V3<R>(orig[3 * k], orig[3 * k + 1], orig[3 * k + 2]), V3<R>(dir[3 * k], dir[3 * k + 1], dir[3 * k + 2])};
    typename TriMesh<R>::Hit hit;
    uint32_t v = 0;
    bool ok = h->mesh.cast(m, ray, max_toi, &hit, &v);
    if (visits) visits[k] = v;
    if (ok) {
      toi[k] = hit.toi; feature[k] = hit.feature;
      for (int d = 0; d < 3; ++d) nrm[3 * k + d] = hit.normal[d];
    } else {
      toi[k] = -1; feature[k] = 0xffffffffu;
      for (int d = 0; d < 3; ++d) nrm[3 * k + d] = 0;
    }
  });
  return 0;
}

// Brute force over every triangle: (toi, face id) lexicographic minimum.
template <class R>
int trimesh_brute(const R* pts, const uint32_t* idx, int64_t ntri, const R* pose7, const R* orig, const R* dir,
                  int64_t n, R max_toi, R* toi, R* nrm, uint32_t* feature, uint32_t* nhits, int nthreads) {
  Iso<R> m = pose_at(pose7, 0);
  parallel_for(n, nthreads, [&](int64_t k) {
    Ray<R> wr{V3<R>(orig[3 * k], orig[3 * k + 1], orig[3 * k + 2]), V3<R>(dir[3 * k], dir[3 * k + 1], dir[3 * k + 2])};
    Ray<R> ray{inv_xform_pt(m, wr.origin), inv_xform_vec(m, wr.dir)};
    bool found = false;
    R bt = 0; V3<R> bn; uint32_t bf = 0; uint32_t cnt = 0;
    for (int64_t f = 0; f < ntri; ++f) {
      const R* a = pts + 3 * idx[3 * f];
      const R* b = pts + 3 * idx[3 * f + 1];
      const R* c = pts + 3 * idx[3 * f + 2];
      R t; V3<R> nn; int fid;
      if (ray_triangle(V3<R>(a[0], a[1], a[2]), V3<R>(b[0], b[1], b[2]), V3<R>(c[0], c[1], c[2]), ray, &t, &nn, &fid) && t <= max_toi) {
        cnt++;
        if (!found || t < bt) { found = true; bt = t; bn = nn; bf = fid == 1 ? (uint32_t)(f + ntri) : (uint32_t)f; }
      }
    }
    if (nhits) nhits[k] = cnt;
    if (found) {
      bn = xform_vec(m, bn);
      toi[k] = bt; feature[k] = bf;
      for (int d = 0; d < 3; ++d) nrm[3 * k + d] = bn[d];
    } else {
      toi[k] = -1; feature[k] = 0xffffffffu;
      for (int d = 0; d < 3; ++d) nrm[3 * k + d] = 0;
    }
  });
  return 0;
}

template <class R> inline AABB<R> aabb_from6(const R* b) {
  return AABB<R>{V3<R>(b[0], b[1], b[2]), V3<R>(b[3], b[4], b[5])};
}

// Per-interaction state of the contact-manifold generators, kept across steps by slot.
struct ManifoldStore {
  std::vector<ManifoldState<float>> slots;
};

int manifold_step(ManifoldStore* st, int64_t ns, const uint32_t* types, const float* params, const uint32_t* hoff,
                  const uint32_t* hcnt, const float* hpts, int64_t npts_total, const uint32_t* hull_tris,
                  const uint32_t* tri_off, const uint32_t* tri_cnt, const float* poses, const uint32_t* ia,
                  const uint32_t* ib, const uint32_t* slots, int64_t np, float linear, float ang1, float ang2,
                  int64_t maxc, uint32_t* count, float* w1, float* w2, float* nrm, float* depth, uint32_t* f1,
                  uint32_t* f2, float* dirs_out, int nthreads) {
  ShapeSet<float> ss;
  ss.load(ns, types, params, hoff, hcnt, hpts, npts_total);
  std::vector<HullPolyO<float>> polys((size_t)ns);
  std::vector<uint8_t> bad((size_t)ns, 0), used((size_t)ns, 0);
  for (int64_t k = 0; k < np; ++k) {
    if (ia[k] >= ns || ib[k] >= ns) return -4;
    used[ia[k]] = used[ib[k]] = 1;
  }
  parallel_for(ns, nthreads, [&](int64_t i) {  // the hulls the pairs name
    Shape<float>& g = ss.shapes[i];
    if (g.type != HULL || !used[i]) return;
    if (!hull_poly_build_o(g.pts, g.npts, hull_tris + 3 * (size_t)tri_off[i], tri_cnt[i], polys[i])) bad[i] = 1;
    g.poly = &polys[i];
  });
  for (int64_t i = 0; i < ns; ++i)
    if (bad[i]) return -1;
  for (int64_t k = 0; k < np; ++k)
    if (slots[k] >= st->slots.size()) return -3;
  const float c1 = std::cos(ang1), s1 = std::sin(ang1), c2 = std::cos(ang2), s2 = std::sin(ang2);
  std::vector<uint8_t> over((size_t)np, 0);
  parallel_for(np, nthreads, [&](int64_t k) {
    uint32_t a = ia[k], b = ib[k];
    ManifoldState<float>& ms = st->slots[slots[k]];
    std::vector<TrackedEntry<float>> cache;
    manifold_update(pose_at(poses, a), ss.shapes[a], pose_at(poses, b), ss.shapes[b], linear, c1, s1, c2, s2, ms, cache);
    int64_t j = 0;
    for (auto& e : cache) {
      if (!e.active) continue;
      if (j < maxc) {
        size_t o = (size_t)k * maxc + j;
        for (int d = 0; d < 3; ++d) {
          w1[3 * o + d] = e.c.world1[d];
          w2[3 * o + d] = e.c.world2[d];
          nrm[3 * o + d] = e.c.normal[d];
        }
        depth[o] = e.c.depth;
        f1[o] = e.f1;
        f2[o] = e.f2;
      }
      j++;
    }
    count[k] = (uint32_t)j;
    if (j > maxc) over[k] = 1;
    for (int d = 0; d < 3; ++d) dirs_out[3 * k + d] = ms.has_dir ? ms.dir[d] : 0.f;
  });
  for (int64_t k = 0; k < np; ++k)
    if (over[k]) return -2;
  return 0;
}

}  // namespace



template <class R>
int toi_batch(int64_t ns, const uint32_t* types, const R* params, const uint32_t* hoff, const uint32_t* hcnt,
              const R* hpts, int64_t npts_total, const R* poses, const R* vel, const uint32_t* ia, const uint32_t* ib,
              int64_t np, R max_toi, R target, uint8_t* has, R* toi, R* w1, R* w2, R* n1, R* n2, uint8_t* status,
              int32_t* iters, int nthreads) {
  ShapeSet<R> ss;
  ss.load(ns, types, params, hoff, hcnt, hpts, npts_total);
  parallel_for(np, nthreads, [&](int64_t k) {
    uint32_t a = ia[k], b = ib[k];
    TOIResult<R> r;
    int it = 0;
    V3<R> v1(vel[3 * a], vel[3 * a + 1], vel[3 * a + 2]), v2(vel[3 * b], vel[3 * b + 1], vel[3 * b + 2]);
    bool ok = time_of_impact(pose_at(poses, a), v1, ss.shapes[a], pose_at(poses, b), v2, ss.shapes[b], max_toi, target,
                             &r, &it);
    has[k] = ok ? 1 : 0;
    if (iters) iters[k] = it;
    if (!ok) { r.toi = 0; r.witness1 = r.witness2 = r.normal1 = r.normal2 = V3<R>(0, 0, 0); r.status = 0; }
    toi[k] = r.toi;
    status[k] = (uint8_t)r.status;
    for (int d = 0; d < 3; ++d) {
      w1[3 * k + d] = r.witness1[d]; w2[3 * k + d] = r.witness2[d];
      n1[3 * k + d] = r.normal1[d]; n2[3 * k + d] = r.normal2[d];
    }
  });
  return 0;
}

template <class R>
int shape_rays(int64_t ns, const uint32_t* types, const R* params, const uint32_t* hoff, const uint32_t* hcnt,
               const R* hpts, int64_t npts_total, const R* poses, const uint32_t* idx, const R* orig, const R* dir,
               int64_t n, R max_toi, int solid, R* toi, R* nrm, uint32_t* feat, int32_t* iters, int nthreads) {
  ShapeSet<R> ss;
  ss.load(ns, types, params, hoff, hcnt, hpts, npts_total);
  parallel_for(n, nthreads, [&](int64_t k) {
    Ray<R> ray{V3<R>(orig[3 * k], orig[3 * k + 1], orig[3 * k + 2]), V3<R>(dir[3 * k], dir[3 * k + 1], dir[3 * k + 2])};
    RayHit<R> h;
    uint32_t i = idx[k];
    int it = 0;
    const bool hit = shape_ray_cast(ss.shapes[i], pose_at(poses, i), ray, max_toi, solid != 0, &h, &it);
    if (iters) iters[k] = it;
    if (hit) {
      toi[k] = h.toi;
      for (int d = 0; d < 3; ++d) nrm[3 * k + d] = h.normal[d];
      feat[k] = h.feature;
    } else {
      toi[k] = R(-1);
      for (int d = 0; d < 3; ++d) nrm[3 * k + d] = R(0);
      feat[k] = 0xFFFFFFFFu;
    }
  });
  return 0;
}


template <class R>
int64_t world_rays(int64_t ns, const uint32_t* types, const R* params, const uint32_t* hoff, const uint32_t* hcnt,
                   const R* hpts, int64_t npts_total, const R* poses, const R* boxes6, const R* orig, const R* dir,
                   int64_t nr, R max_toi, int mode, uint32_t* handles, R* toi, R* nrm, int64_t* counts, int64_t cap,
                   int nthreads) {
  ShapeSet<R> ss;
  ss.load(ns, types, params, hoff, hcnt, hpts, npts_total);
  std::vector<Iso<R>> ps((size_t)ns);
  std::vector<AABB<R>> bb((size_t)ns);
  for (int64_t i = 0; i < ns; ++i) {
    ps[i] = pose_at(poses, i);
    bb[i] = aabb_from6(boxes6 + 6 * i);
  }
  auto ray_at = [&](int64_t k) {
    return Ray<R>{V3<R>(orig[3 * k], orig[3 * k + 1], orig[3 * k + 2]), V3<R>(dir[3 * k], dir[3 * k + 1], dir[3 * k + 2])};
  };
  if (mode == 0) {  // first interference
    parallel_for(nr, nthreads, [&](int64_t k) {
      uint32_t h = 0xFFFFFFFFu;
      RayHit<R> r;
      if (world_first_ray(ss.shapes, ps, bb.data(), ray_at(k), max_toi, &h, &r)) {
        handles[k] = h; toi[k] = r.toi;
        for (int d = 0; d < 3; ++d) nrm[3 * k + d] = r.normal[d];
      } else {
        handles[k] = 0xFFFFFFFFu; toi[k] = R(-1);
        for (int d = 0; d < 3; ++d) nrm[3 * k + d] = R(0);
      }
    });
    return nr;
  }
  // mode 1: interferences_with_ray with shape casts; mode 2: stored boxes only.  Handles ascending.
  std::vector<std::vector<uint32_t>> hh((size_t)nr);
  std::vector<std::vector<RayHit<R>>> rr((size_t)nr);
  parallel_for(nr, nthreads, [&](int64_t k) {
    Ray<R> ray = ray_at(k);
    for (int64_t h = 0; h < ns; ++h) {
      R bt;
      if (!ray_aabb_toi(bb[h], ray, max_toi, true, &bt)) continue;
      RayHit<R> r{R(0), V3<R>(0, 0, 0), 0};
      if (mode == 1 && !shape_ray_cast(ss.shapes[h], ps[h], ray, max_toi, true, &r)) continue;
      hh[k].push_back((uint32_t)h);
      rr[k].push_back(r);
    }
  });
  int64_t total = 0;
  for (int64_t k = 0; k < nr; ++k) { counts[k] = (int64_t)hh[k].size(); total += counts[k]; }
  if (total <= cap) {
    int64_t o = 0;
    for (int64_t k = 0; k < nr; ++k)
      for (size_t j = 0; j < hh[k].size(); ++j, ++o) {
        handles[o] = hh[k][j]; toi[o] = rr[k][j].toi;
        for (int d = 0; d < 3; ++d) nrm[3 * o + d] = rr[k][j].normal[d];
      }
  }
  return total;
}


template <class R>
int polyhedra_step_batch(int64_t ns, const uint32_t* types, const R* params, const uint32_t* hoff, const uint32_t* hcnt,
                         const R* hpts, int64_t npts_total, const R* poses, const uint32_t* ia, const uint32_t* ib,
                         int64_t np, R prediction, const R* last_dirs, uint8_t* has, R* w1, R* w2, R* nrm, R* depth,
                         R* dir_out, uint8_t* dir_set, int nthreads) {
  ShapeSet<R> ss;
  ss.load(ns, types, params, hoff, hcnt, hpts, npts_total);
  parallel_for(np, nthreads, [&](int64_t k) {
    uint32_t a = ia[k], b = ib[k];
    V3<R> ld(last_dirs[3 * k], last_dirs[3 * k + 1], last_dirs[3 * k + 2]);
    const bool some = ld.x != R(0) || ld.y != R(0) || ld.z != R(0);
    Contact<R> c;
    V3<R> nd(0, 0, 0);
    bool set = false;
    bool ok = polyhedra_gjk_step(pose_at(poses, a), ss.shapes[a], pose_at(poses, b), ss.shapes[b], prediction,
                                 some ? &ld : nullptr, &c, &nd, &set);
    has[k] = ok ? 1 : 0;
    dir_set[k] = set ? 1 : 0;
    for (int d = 0; d < 3; ++d) {
      w1[3 * k + d] = ok ? c.world1[d] : R(0);
      w2[3 * k + d] = ok ? c.world2[d] : R(0);
      nrm[3 * k + d] = ok ? c.normal[d] : R(0);
      dir_out[3 * k + d] = nd[d];
    }
    depth[k] = ok ? c.depth : R(0);
  });
  return 0;
}


template <class R>
int mesh_queries(const R* tri_pts, int64_t ntri, const R* mesh_pose7, int64_t ns, const uint32_t* types, const R* params,
                 const uint32_t* hoff, const uint32_t* hcnt, const R* hpts, int64_t npts_total, const R* poses,
                 const uint32_t* idx, int64_t nq, R prediction, int mode, uint8_t* has, R* w1, R* w2, R* nrm, R* depth,
                 uint32_t* face, R* dist, int nthreads) {
  ShapeSet<R> ss;
  ss.load(ns, types, params, hoff, hcnt, hpts, npts_total);
  std::vector<V3<R>> tri((size_t)ntri * 3);
  for (int64_t i = 0; i < 3 * ntri; ++i) tri[i] = V3<R>(tri_pts[3 * i], tri_pts[3 * i + 1], tri_pts[3 * i + 2]);
  Iso<R> m1 = pose_at(mesh_pose7, 0);
  parallel_for(nq, nthreads, [&](int64_t k) {
    uint32_t s = idx[k];
    if (mode == 0) {
      Contact<R> c;
      c.depth = R(0);
      uint32_t f = 0xFFFFFFFFu;
      bool ok = contact_mesh_shape(m1, tri.data(), (size_t)ntri, pose_at(poses, s), ss.shapes[s], prediction, &c, &f);
      has[k] = ok ? 1 : 0;
      face[k] = f;
      for (int d = 0; d < 3; ++d) {
        w1[3 * k + d] = ok ? c.world1[d] : R(0);
        w2[3 * k + d] = ok ? c.world2[d] : R(0);
        nrm[3 * k + d] = ok ? c.normal[d] : R(0);
      }
      depth[k] = ok ? c.depth : R(0);
    } else {
      dist[k] = distance_mesh_shape(m1, tri.data(), (size_t)ntri, pose_at(poses, s), ss.shapes[s]);
    }
  });
  return 0;
}

extern "C" {

int orc_contact_f32(int64_t ns, const uint32_t* types, const float* params, const uint32_t* hoff, const uint32_t* hcnt,
                    const float* hpts, int64_t npts, const float* poses, const uint32_t* ia, const uint32_t* ib,
                    int64_t np, float prediction, uint8_t* has, float* w1, float* w2, float* n, float* depth,
                    int32_t* gi, int32_t* ei, int nthreads) {
  return contact_batch<float>(ns, types, params, hoff, hcnt, hpts, npts, poses, ia, ib, np, prediction, has, w1, w2, n,
                              depth, gi, ei, nthreads);
}
int orc_contact_f64(int64_t ns, const uint32_t* types, const double* params, const uint32_t* hoff,
                    const uint32_t* hcnt, const double* hpts, int64_t npts, const double* poses, const uint32_t* ia,
                    const uint32_t* ib, int64_t np, double prediction, uint8_t* has, double* w1, double* w2,
                    double* n, double* depth, int32_t* gi, int32_t* ei, int nthreads) {
  return contact_batch<double>(ns, types, params, hoff, hcnt, hpts, npts, poses, ia, ib, np, prediction, has, w1, w2,
                               n, depth, gi, ei, nthreads);
}
int orc_contact_sm_f32(int64_t ns, const uint32_t* types, const float* params, const uint32_t* hoff,
                       const uint32_t* hcnt, const float* hpts, int64_t npts, const float* poses, const uint32_t* ia,
                       const uint32_t* ib, int64_t np, float prediction, const float* init_dirs, int32_t max_iter,
                       uint8_t* has, float* w1, float* w2, float* n, float* depth, int32_t* gi, int32_t* ei,
                       int32_t* st, int nthreads) {
  return contact_sm_batch<float>(ns, types, params, hoff, hcnt, hpts, npts, poses, ia, ib, np, prediction, init_dirs,
                                 max_iter, has, w1, w2, n, depth, gi, ei, st, nthreads);
}
int orc_contact_sm_f64(int64_t ns, const uint32_t* types, const double* params, const uint32_t* hoff,
                       const uint32_t* hcnt, const double* hpts, int64_t npts, const double* poses, const uint32_t* ia,
                       const uint32_t* ib, int64_t np, double prediction, const double* init_dirs, int32_t max_iter,
                       uint8_t* has, double* w1, double* w2, double* n, double* depth, int32_t* gi, int32_t* ei,
                       int32_t* st, int nthreads) {
  return contact_sm_batch<double>(ns, types, params, hoff, hcnt, hpts, npts, poses, ia, ib, np, prediction, init_dirs,
                                  max_iter, has, w1, w2, n, depth, gi, ei, st, nthreads);
}
int orc_distance_f32(int64_t ns, const uint32_t* types, const float* params, const uint32_t* hoff,
                     const uint32_t* hcnt, const float* hpts, int64_t npts, const float* poses, const uint32_t* ia,
                     const uint32_t* ib, int64_t np, const float* init_dirs, int32_t max_iter, float* dist, int32_t* gi,
                     int nthreads) {
  return distance_batch<float>(ns, types, params, hoff, hcnt, hpts, npts, poses, ia, ib, np, init_dirs, max_iter, dist,
                               gi, nthreads);
}
int orc_distance_f64(int64_t ns, const uint32_t* types, const double* params, const uint32_t* hoff,
                     const uint32_t* hcnt, const double* hpts, int64_t npts, const double* poses, const uint32_t* ia,
                     const uint32_t* ib, int64_t np, const double* init_dirs, int32_t max_iter, double* dist,
                     int32_t* gi, int nthreads) {
  return distance_batch<double>(ns, types, params, hoff, hcnt, hpts, npts, poses, ia, ib, np, init_dirs, max_iter,
                                dist, gi, nthreads);
}
int orc_shape_aabb_f32(int64_t ns, const uint32_t* types, const float* params, const uint32_t* hoff,
                       const uint32_t* hcnt, const float* hpts, int64_t npts, const float* poses, float* out6) {
  return shape_aabbs<float>(ns, types, params, hoff, hcnt, hpts, npts, poses, out6);
}
int orc_shape_aabb_f64(int64_t ns, const uint32_t* types, const double* params, const uint32_t* hoff,
                       const uint32_t* hcnt, const double* hpts, int64_t npts, const double* poses, double* out6) {
  return shape_aabbs<double>(ns, types, params, hoff, hcnt, hpts, npts, poses, out6);
}

// ---- TriMesh
void* orc_trimesh_create_f32(const float* pts, int64_t npts, const uint32_t* idx, int64_t ntri) {
  return trimesh_create<float>(pts, npts, idx, ntri);
}
void orc_trimesh_free_f32(void* h) { delete (MeshHandle<float>*)h; }
int orc_trimesh_cast_f32(void* h, const float* pose7, const float* orig, const float* dir, int64_t n, float max_toi,
                         float* toi, float* nrm, uint32_t* feature, uint32_t* visits, int nthreads) {
  return trimesh_cast<float>(h, pose7, orig, dir, n, max_toi, toi, nrm, feature, visits, nthreads);
}
void* orc_trimesh_create_f64(const double* pts, int64_t npts, const uint32_t* idx, int64_t ntri) {
  return trimesh_create<double>(pts, npts, idx, ntri);
}
void orc_trimesh_free_f64(void* h) { delete (MeshHandle<double>*)h; }
int orc_trimesh_cast_f64(void* h, const double* pose7, const double* orig, const double* dir, int64_t n,
                         double max_toi, double* toi, double* nrm, uint32_t* feature, uint32_t* visits, int nthreads) {
  return trimesh_cast<double>(h, pose7, orig, dir, n, max_toi, toi, nrm, feature, visits, nthreads);
}
int orc_trimesh_brute_f32(const float* pts, const uint32_t* idx, int64_t ntri, const float* pose7, const float* orig,
                          const float* dir, int64_t n, float max_toi, float* toi, float* nrm, uint32_t* feature,
                          uint32_t* nhits, int nthreads) {
  return trimesh_brute<float>(pts, idx, ntri, pose7, orig, dir, n, max_toi, toi, nrm, feature, nhits, nthreads);
}
int orc_trimesh_brute_f64(const double* pts, const uint32_t* idx, int64_t ntri, const double* pose7,
                          const double* orig, const double* dir, int64_t n, double max_toi, double* toi, double* nrm,
                          uint32_t* feature, uint32_t* nhits, int nthreads) {
  return trimesh_brute<double>(pts, idx, ntri, pose7, orig, dir, n, max_toi, toi, nrm, feature, nhits, nthreads);
}
// Single ray vs triangle (for closed-form tests).
int orc_ray_triangle_f64(const double* abc, const double* orig, const double* dir, double* toi, double* n, int* fid) {
  Ray<double> r{V3<double>(orig[0], orig[1], orig[2]), V3<double>(dir[0], dir[1], dir[2])};
  V3<double> nn;
  bool ok = ray_triangle(V3<double>(abc[0], abc[1], abc[2]), V3<double>(abc[3], abc[4], abc[5]),
                         V3<double>(abc[6], abc[7], abc[8]), r, toi, &nn, fid);
  for (int d = 0; d < 3; ++d) n[d] = nn[d];
  return ok ? 1 : 0;
}
int orc_ray_aabb_f64(const double* bb6, const double* orig, const double* dir, double max_toi, int solid, double* toi) {
  Ray<double> r{V3<double>(orig[0], orig[1], orig[2]), V3<double>(dir[0], dir[1], dir[2])};
  return ray_aabb_toi(aabb_from6(bb6), r, max_toi, solid != 0, toi) ? 1 : 0;
}

// ---- Voronoi projections of the origin (solid), for brute-force pinning.
int orc_project_f64(int kind, const double* pts, double* proj, int* loc_kind, int* loc_id, double* bc) {
  Loc<double> loc{0, 0, {0, 0, 0}};
  V3<double> P[4];
  for (int i = 0; i < 4; ++i) P[i] = V3<double>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  V3<double> o(0, 0, 0), r;
  if (kind == 1) r = project_segment(P[0], P[1], o, &loc);
  else if (kind == 2) r = project_triangle(P[0], P[1], P[2], o, true, &loc);
  else r = project_tetrahedron(P[0], P[1], P[2], P[3], o, &loc);
  for (int d = 0; d < 3; ++d) proj[d] = r[d];
  *loc_kind = loc.kind; *loc_id = loc.id;
  for (int d = 0; d < 3; ++d) bc[d] = loc.bc[d];
  return 0;
}

// ---- single-function probes for the property tests
// orthonormal_subspace_first (EPA's third point direction on a 1-simplex); returns the vector.
void orc_orthonormal_first_f64(const double* v, double* out) {
  V3<double> r = orthonormal_subspace_first(V3<double>(v[0], v[1], v[2]));
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
int orc_affinely_dependent_f64(const double* p9) {
  return is_affinely_dependent_triangle(V3<double>(p9[0], p9[1], p9[2]), V3<double>(p9[3], p9[4], p9[5]),
                                        V3<double>(p9[6], p9[7], p9[8])) ? 1 : 0;
}
int orc_affinely_dependent_f32(const float* p9) {
  return is_affinely_dependent_triangle(V3<float>(p9[0], p9[1], p9[2]), V3<float>(p9[3], p9[4], p9[5]),
                                        V3<float>(p9[6], p9[7], p9[8])) ? 1 : 0;
}
// ---- DBVTBroadPhase
void* orc_bp_create(double margin) { return new DBVTBroadPhase<float>((float)margin); }
void orc_bp_free(void* h) { delete (DBVTBroadPhase<float>*)h; }
uint32_t orc_bp_create_proxy(void* h, const float* bb6) { return ((DBVTBroadPhase<float>*)h)->create_proxy(aabb_from6(bb6)); }
void orc_bp_set_bv(void* h, uint32_t handle, const float* bb6) {
  ((DBVTBroadPhase<float>*)h)->deferred_set_bounding_volume(handle, aabb_from6(bb6));
}
void orc_bp_remove(void* h, const uint32_t* hs, int64_t n) {
  std::vector<uint32_t> v(hs, hs + n);
  ((DBVTBroadPhase<float>*)h)->remove(v);
}
void orc_bp_update(void* h) { ((DBVTBroadPhase<float>*)h)->update(); }
// CollisionGroups per handle for the pair filter: can_interact_with_groups
// (pipeline/object/collision_groups.rs:353-359); handles never set keep CollisionGroups::new().
void orc_bp_set_groups(void* h, uint32_t handle, uint32_t membership, uint32_t whitelist, uint32_t blacklist) {
  auto* bp = (DBVTBroadPhase<float>*)h;
  static thread_local std::vector<std::pair<void*, std::shared_ptr<std::vector<uint32_t>>>> tables;
  std::shared_ptr<std::vector<uint32_t>> t;
  for (auto& e : tables) if (e.first == h) t = e.second;
  if (!t) {
    t = std::make_shared<std::vector<uint32_t>>();
    tables.push_back({h, t});
    bp->allowed = [t](uint32_t a, uint32_t b) {
      const uint32_t all = (1u << 30) - 1u;
      auto get = [&](uint32_t x, int k) { return (size_t)3 * x + k < t->size() ? (*t)[3 * x + k] : (k == 2 ? 0u : all); };
      uint32_t ma = get(a, 0), wa = get(a, 1), ba = get(a, 2), mb = get(b, 0), wb = get(b, 1), bb = get(b, 2);
      return (ma & bb) == 0 && (mb & ba) == 0 && (ma & wb) != 0 && (mb & wa) != 0;
    };
  }
  const uint32_t all = (1u << 30) - 1u;
  while (t->size() < (size_t)3 * (handle + 1)) { t->push_back(all); t->push_back(all); t->push_back(0u); }
  (*t)[3 * handle] = membership; (*t)[3 * handle + 1] = whitelist; (*t)[3 * handle + 2] = blacklist;
}
int64_t orc_bp_num_pairs(void* h) { return (int64_t)((DBVTBroadPhase<float>*)h)->pairs.size(); }
// Sorted (a << 32 | b) keys with a < b.
int64_t orc_bp_pairs(void* h, uint64_t* out, int64_t cap) {
  auto* bp = (DBVTBroadPhase<float>*)h;
  std::vector<uint64_t> k;
  for (auto& p : bp->pairs) k.push_back(p.first);
  std::sort(k.begin(), k.end());
  int64_t n = std::min<int64_t>(cap, (int64_t)k.size());
  for (int64_t i = 0; i < n; ++i) out[i] = k[i];
  return (int64_t)k.size();
}
int64_t orc_bp_events(void* h, int which, uint64_t* out, int64_t cap) {
  auto* bp = (DBVTBroadPhase<float>*)h;
  std::vector<uint64_t> k = which == 0 ? bp->started : bp->stopped;
  std::sort(k.begin(), k.end());
  int64_t n = std::min<int64_t>(cap, (int64_t)k.size());
  for (int64_t i = 0; i < n; ++i) out[i] = k[i];
  return (int64_t)k.size();
}
void orc_bp_clear_events(void* h) {
  auto* bp = (DBVTBroadPhase<float>*)h;
  bp->started.clear(); bp->stopped.clear();
}
int orc_bp_stored_bv(void* h, uint32_t handle, float* out6) {
  auto* bp = (DBVTBroadPhase<float>*)h;
  const AABB<float>* b = bp->stored_bv(handle);
  if (!b) return 0;
  out6[0] = b->mins.x; out6[1] = b->mins.y; out6[2] = b->mins.z;
  out6[3] = b->maxs.x; out6[4] = b->maxs.y; out6[5] = b->maxs.z;
  return 1;
}
// O(N^2) brute-force pair set over AoS boxes (6 floats each): sorted keys i<j.
int64_t orc_pairs_brute(const float* bb6, int64_t n, uint64_t* out, int64_t cap) {
  int64_t cnt = 0;
  for (int64_t i = 0; i < n; ++i) {
    AABB<float> a = aabb_from6(bb6 + 6 * i);
    for (int64_t j = i + 1; j < n; ++j) {
      if (aabb_intersects(a, aabb_from6(bb6 + 6 * j))) {
        if (cnt < cap) out[cnt] = ((uint64_t)i << 32) | (uint64_t)j;
        cnt++;
      }
    }
  }
  return cnt;
}

#define ORC_TOI(T, R)                                                                                             \
  int orc_toi_##T(int64_t ns, const uint32_t* types, const R* params, const uint32_t* hoff, const uint32_t* hcnt,  \
                  const R* hpts, int64_t npts, const R* poses, const R* vel, const uint32_t* ia, const uint32_t* ib, \
                  int64_t np, R max_toi, R target, uint8_t* has, R* toi, R* w1, R* w2, R* n1, R* n2,               \
                  uint8_t* status, int32_t* iters, int nthreads) {                                                 \
    return toi_batch<R>(ns, types, params, hoff, hcnt, hpts, npts, poses, vel, ia, ib, np, max_toi, target, has,   \
                        toi, w1, w2, n1, n2, status, iters, nthreads);                                             \
  }                                                                                                              \
  int orc_shape_rays_##T(int64_t ns, const uint32_t* types, const R* params, const uint32_t* hoff,                \
                         const uint32_t* hcnt, const R* hpts, int64_t npts, const R* poses, const uint32_t* idx,   \
                         const R* orig, const R* dir, int64_t n, R max_toi, int solid, R* toi, R* nrm,              \
                         uint32_t* feat, int32_t* iters, int nthreads) {                                           \
    return shape_rays<R>(ns, types, params, hoff, hcnt, hpts, npts, poses, idx, orig, dir, n, max_toi, solid, toi, \
                         nrm, feat, iters, nthreads);                                                              \
  }
ORC_TOI(f32, float)
ORC_TOI(f64, double)

int64_t orc_world_rays_f32(int64_t ns, const uint32_t* types, const float* params, const uint32_t* hoff,
                           const uint32_t* hcnt, const float* hpts, int64_t npts, const float* poses, const float* boxes6,
                           const float* orig, const float* dir, int64_t nr, float max_toi, int mode, uint32_t* handles,
                           float* toi, float* nrm, int64_t* counts, int64_t cap, int nthreads) {
  return world_rays<float>(ns, types, params, hoff, hcnt, hpts, npts, poses, boxes6, orig, dir, nr, max_toi, mode,
                           handles, toi, nrm, counts, cap, nthreads);
}

int orc_polyhedra_step_f32(int64_t ns, const uint32_t* types, const float* params, const uint32_t* hoff,
                           const uint32_t* hcnt, const float* hpts, int64_t npts, const float* poses, const uint32_t* ia,
                           const uint32_t* ib, int64_t np, float prediction, const float* last_dirs, uint8_t* has,
                           float* w1, float* w2, float* nrm, float* depth, float* dir_out, uint8_t* dir_set,
                           int nthreads) {
  return polyhedra_step_batch<float>(ns, types, params, hoff, hcnt, hpts, npts, poses, ia, ib, np, prediction,
                                     last_dirs, has, w1, w2, nrm, depth, dir_out, dir_set, nthreads);
}

#define ORC_MESH(T, R)                                                                                           \
  int orc_mesh_queries_##T(const R* tri_pts, int64_t ntri, const R* mesh_pose7, int64_t ns, const uint32_t* types, \
                           const R* params, const uint32_t* hoff, const uint32_t* hcnt, const R* hpts, int64_t npts,  \
                           const R* poses, const uint32_t* idx, int64_t nq, R prediction, int mode, uint8_t* has,     \
                           R* w1, R* w2, R* nrm, R* depth, uint32_t* face, R* dist, int nthreads) {                  \
    return mesh_queries<R>(tri_pts, ntri, mesh_pose7, ns, types, params, hoff, hcnt, hpts, npts, poses, idx, nq,       \
                           prediction, mode, has, w1, w2, nrm, depth, face, dist, nthreads);                          \
  }
ORC_MESH(f32, float)
ORC_MESH(f64, double)


void* orc_manifold_store_new(int64_t nslots) {
  ManifoldStore* st = new ManifoldStore;
  st->slots.resize((size_t)nslots);
  return st;
}
void orc_manifold_store_free(void* p) { delete (ManifoldStore*)p; }
int orc_manifold_reset(void* p, const uint32_t* slots, int64_t n) {
  ManifoldStore* st = (ManifoldStore*)p;
  for (int64_t k = 0; k < n; ++k) {
    if (slots[k] >= st->slots.size()) return -3;
    st->slots[slots[k]] = ManifoldState<float>();
  }
  return 0;
}
int orc_manifold_step_f32(void* store, int64_t ns, const uint32_t* types, const float* params, const uint32_t* hoff,
                          const uint32_t* hcnt, const float* hpts, int64_t npts, const uint32_t* hull_tris,
                          const uint32_t* tri_off, const uint32_t* tri_cnt, const float* poses, const uint32_t* ia,
                          const uint32_t* ib, const uint32_t* slots, int64_t np, float linear, float ang1, float ang2,
                          int64_t maxc, uint32_t* count, float* w1, float* w2, float* nrm, float* depth, uint32_t* f1,
                          uint32_t* f2, float* dirs_out, int nthreads) {
  return manifold_step((ManifoldStore*)store, ns, types, params, hoff, hcnt, hpts, npts, hull_tris, tri_off, tri_cnt,
                       poses, ia, ib, slots, np, linear, ang1, ang2, maxc, count, w1, w2, nrm, depth, f1, f2, dirs_out,
                       nthreads);
}
int orc_hull_polyhedron_f32(const float* pts, int64_t npts, const uint32_t* tris, int64_t ntri, float* face_normals,
                            uint32_t* face_first, uint32_t* face_count, int64_t* nfaces, uint32_t* loop_vertex,
                            uint32_t* loop_edge, int64_t* nloops, uint32_t* edge_vertices, float* edge_dirs,
                            int64_t* nedges, uint32_t* vertex_first, uint32_t* vertex_count, uint32_t* adj_face,
                            uint32_t* adj_edge) {
  std::vector<V3<float>> P((size_t)npts);
  for (int64_t i = 0; i < npts; ++i) P[i] = V3<float>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  HullPolyO<float> H;
  if (!hull_poly_build_o(P.data(), (uint32_t)npts, tris, (uint32_t)ntri, H)) return -1;
  *nfaces = (int64_t)H.ffirst.size();
  *nloops = (int64_t)H.lvert.size();
  *nedges = (int64_t)H.ev0.size();
  for (size_t f = 0; f < H.ffirst.size(); ++f) {
    for (int d = 0; d < 3; ++d) face_normals[3 * f + d] = H.fn[f][d];
    face_first[f] = H.ffirst[f];
    face_count[f] = H.fnum[f];
  }
  for (size_t k = 0; k < H.lvert.size(); ++k) { loop_vertex[k] = H.lvert[k]; loop_edge[k] = H.ledge[k]; }
  for (size_t e = 0; e < H.ev0.size(); ++e) {
    edge_vertices[2 * e] = H.ev0[e];
    edge_vertices[2 * e + 1] = H.ev1[e];
    for (int d = 0; d < 3; ++d) edge_dirs[3 * e + d] = H.edir[e][d];
  }
  for (int64_t v = 0; v < npts; ++v) { vertex_first[v] = H.vfirst[v]; vertex_count[v] = H.vnum[v]; }
  for (size_t k = 0; k < H.vface.size(); ++k) { adj_face[k] = H.vface[k]; adj_edge[k] = H.vedge[k]; }
  return 0;
}

}  // extern "C"
