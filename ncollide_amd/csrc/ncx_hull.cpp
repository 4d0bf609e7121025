// ncx_hull.cpp -- convex hull vertex selection and polyhedral structure (input preparation,
// host code).
//
// Stands in for ConvexHull::try_from_points -> transformation::convex_hull3
// (src/shape/convex.rs:88-97, src/transformation/convex_hull3.rs:28-149).  The
// reference runs quickhull on PCA-whitened points and maps the result back,
// which perturbs vertex coordinates by rounding and reorders them; here the
// hull is built incrementally in double precision and the result is the list of
// input indices that are hull vertices, in input order, so hull vertices are
// bit-identical to the input points.  Support functions only read the vertex
// list; the contact-manifold generators also read the faces, edges and adjacency
// that ConvexHull::try_new derives from the hull's triangles (hull_poly_build).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <thread>
#include <vector>
#include "../../include/ncx.h"
#include "ncx_hullpoly.h"

namespace {

struct D3 {
  double x, y, z;
};
inline D3 sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 crs(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dt(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double nrm(D3 a) { return std::sqrt(dt(a, a)); }

struct Face {
  int v[3];
  D3 n;  // outward, unnormalised
  double off;
  bool alive;
};

struct HullBuilder {
  std::vector<Face> faces;
  // directed edge (a, b) -> the face holding it: per start vertex a, a short list of (b, face);
  // O(edges) memory whatever the point count
  std::vector<std::vector<std::pair<int, int>>> adj;
  int n = 0;
  std::vector<int> visible;
  std::vector<std::pair<int, int>> horizon;
  std::vector<uint8_t> is_vis;

  int get_edge(int a, int b) const {
    for (const auto& e : adj[a])
      if (e.first == b) return e.second;
    return -1;
  }
  void set_edge(int a, int b, int f) {
    for (auto& e : adj[a])
      if (e.first == b) {
        e.second = f;
        return;
      }
    adj[a].push_back({b, f});
  }
  void add_face(const std::vector<D3>& P, int a, int b, int c) {
    Face f;
    f.v[0] = a; f.v[1] = b; f.v[2] = c;
    f.n = crs(sub(P[b], P[a]), sub(P[c], P[a]));
    f.off = dt(f.n, P[a]);
    f.alive = true;
    int id = (int)faces.size();
    faces.push_back(f);
    set_edge(a, b, id);
    set_edge(b, c, id);
    set_edge(c, a, id);
  }

  // Builds the hull; false if degenerate.  Alive faces are outward triangles of input indices.
  bool run(const float* xyz, int npts) {
    faces.clear();
    if (npts < 4) return false;
    n = npts;
    std::vector<D3> P(npts);
    double scale = 0.0;
    for (int i = 0; i < npts; ++i) {
      P[i] = D3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
      scale = std::max(scale, std::max(std::fabs(P[i].x), std::max(std::fabs(P[i].y), std::fabs(P[i].z))));
    }
    if (scale == 0.0) return false;
    const double eps = 1e-11 * scale;
    // Initial tetrahedron: extreme point, farthest point, farthest from the line, farthest from the plane.
    int i0 = 0;
    for (int i = 1; i < npts; ++i) if (P[i].x < P[i0].x) i0 = i;
    int i1 = -1;
    double best = 0;
    for (int i = 0; i < npts; ++i) { double d = nrm(sub(P[i], P[i0])); if (d > best) { best = d; i1 = i; } }
    if (i1 < 0 || best <= eps) return false;
    D3 l = sub(P[i1], P[i0]);
    int i2 = -1;
    best = 0;
    for (int i = 0; i < npts; ++i) { double d = nrm(crs(l, sub(P[i], P[i0]))); if (d > best) { best = d; i2 = i; } }
    if (i2 < 0 || best <= eps * nrm(l)) return false;
    D3 pn = crs(l, sub(P[i2], P[i0]));
    int i3 = -1;
    best = 0;
    for (int i = 0; i < npts; ++i) { double d = std::fabs(dt(pn, sub(P[i], P[i0]))); if (d > best) { best = d; i3 = i; } }
    if (i3 < 0 || best <= eps * nrm(pn)) return false;
    adj.assign(npts, {});
    if (dt(pn, sub(P[i3], P[i0])) > 0) std::swap(i1, i2);  // make (i0, i1, i2) face away from i3
    add_face(P, i0, i1, i2);
    add_face(P, i0, i3, i1);
    add_face(P, i1, i3, i2);
    add_face(P, i2, i3, i0);
    is_vis.assign(4 * npts + 16, 0);
    for (int p = 0; p < npts; ++p) {
      if (p == i0 || p == i1 || p == i2 || p == i3) continue;
      visible.clear();
      for (int f = 0; f < (int)faces.size(); ++f) {
        if (!faces[f].alive) continue;
        double s = dt(faces[f].n, P[p]) - faces[f].off;
        if (s > eps * nrm(faces[f].n)) visible.push_back(f);
      }
      if (visible.empty()) continue;
      if (is_vis.size() < faces.size() + 64) is_vis.resize(faces.size() * 2 + 64, 0);
      for (int f : visible) is_vis[f] = 1;
      horizon.clear();
      for (int f : visible) {
        for (int e = 0; e < 3; ++e) {
          int a = faces[f].v[e], b = faces[f].v[(e + 1) % 3];
          int g = get_edge(b, a);
          if (g < 0 || !is_vis[g]) horizon.push_back({a, b});
        }
      }
      for (int f : visible) {
        faces[f].alive = false;
        is_vis[f] = 0;
      }
      for (auto& e : horizon) add_face(P, e.first, e.second, p);
    }
    return true;
  }

  // Returns the number of hull vertices written to `out` (input indices, ascending), 0 if degenerate.
  int build(const float* xyz, int npts, uint32_t* out) {
    if (!run(xyz, npts)) return 0;
    std::vector<uint8_t> used(npts, 0);
    for (auto& f : faces)
      if (f.alive) used[f.v[0]] = used[f.v[1]] = used[f.v[2]] = 1;
    int k = 0;
    for (int i = 0; i < npts; ++i) if (used[i]) out[k++] = (uint32_t)i;
    return k;
  }
};

// f32 helpers in the reference's operation order (nalgebra Vector3<f32>)
struct F3 {
  float x, y, z;
};
inline F3 fsub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float fdot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline F3 fcross(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// Unit::try_new(v, eps): None when |v| <= eps, else v / |v| (componentwise division)
inline bool ftry_normalize(F3 v, float eps, F3* out) {
  float n = std::sqrt(fdot(v, v));
  if (!(n > eps)) return false;
  *out = F3{v.x / n, v.y / n, v.z / n};
  return true;
}

}  // namespace

bool hull_triangulate(const float* xyz, uint32_t npts, std::vector<uint32_t>& tris) {
  tris.clear();
  HullBuilder hb;
  if (!hb.run(xyz, (int)npts)) return false;
  for (auto& f : hb.faces)
    if (f.alive) {
      tris.push_back((uint32_t)f.v[0]);
      tris.push_back((uint32_t)f.v[1]);
      tris.push_back((uint32_t)f.v[2]);
    }
  return true;
}

// ConvexHull::try_new (convex.rs:107-324), step by step; f32 throughout, as the f32 reference.
bool hull_poly_build(const float* x, const float* y, const float* z, uint32_t npts, const uint32_t* tris,
                     uint32_t ntri, HullPoly& H) {
  H = HullPoly();
  const float feps = 1.1920929e-7f;
  const float eps = std::sqrt(feps);  // N::default_epsilon().sqrt()
  auto P = [&](uint32_t i) { return F3{x[i], y[i], z[i]}; };
  struct Edge {
    uint32_t v[2];
    uint32_t f[2];
    F3 dir;
    bool deleted;
  };
  struct Tri {
    uint32_t v[3];
    uint32_t e[3];
    F3 n;
    int64_t parent;  // -1: None
  };
  std::vector<Edge> edges;
  std::vector<Tri> T(ntri);
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> emap(npts);  // min vertex -> (max vertex, edge)
  for (uint32_t t = 0; t < ntri; ++t) {
    const uint32_t* vtx = tris + 3 * (size_t)t;
    for (int i = 0; i < 3; ++i)
      if (vtx[i] >= npts) return false;
    for (int i1 = 0; i1 < 3; ++i1) {
      const int i2 = (i1 + 1) % 3;
      const uint32_t lo = std::min(vtx[i1], vtx[i2]), hi = std::max(vtx[i1], vtx[i2]);
      int64_t found = -1;
      for (auto& e : emap[lo])
        if (e.first == hi) found = e.second;
      if (found >= 0) {
        T[t].e[i1] = (uint32_t)found;
        edges[found].f[1] = t;
      } else {
        F3 dir;
        if (!ftry_normalize(fsub(P(vtx[i2]), P(vtx[i1])), feps, &dir)) return false;
        T[t].e[i1] = (uint32_t)edges.size();
        emap[lo].push_back({hi, (uint32_t)edges.size()});
        edges.push_back(Edge{{vtx[i1], vtx[i2]}, {t, 0}, dir, false});
      }
    }
    // utils::ccw_face_normal (utils/ccw_face_normal.rs:21-27)
    F3 nn;
    if (!ftry_normalize(fcross(fsub(P(vtx[1]), P(vtx[0])), fsub(P(vtx[2]), P(vtx[0]))), feps, &nn)) return false;
    T[t].v[0] = vtx[0]; T[t].v[1] = vtx[1]; T[t].v[2] = vtx[2];
    T[t].n = nn;
    T[t].parent = -1;
  }
  // edges between nearly coplanar triangles are deleted
  int64_t num_valid_edges = 0;
  for (auto& e : edges) {
    if (fdot(T[e.f[0]].n, T[e.f[1]].n) > 1.0f - eps) e.deleted = true;
    else num_valid_edges++;
  }
  // faces by following contours
  std::vector<uint32_t> faces_tri;  // the triangle whose normal the face takes
  const size_t guard = 8 * (size_t)ntri + 16;
  for (uint32_t i = 0; i < ntri; ++i) {
    if (T[i].parent >= 0) continue;
    for (int j1 = 0; j1 < 3; ++j1) {
      if (edges[T[i].e[j1]].deleted) continue;
      const uint32_t new_face_id = (uint32_t)H.ffirst.size();
      const uint32_t first = (uint32_t)H.ledge.size();
      uint32_t num = 1;
      H.ledge.push_back(T[i].e[j1]);
      H.lvert.push_back(T[i].v[j1]);
      const uint32_t start_vertex = T[i].v[j1];
      uint32_t ct = i;
      int ce = (j1 + 1) % 3;
      size_t steps = 0;
      while (T[ct].v[ce] != start_vertex) {
        if (++steps > guard) return false;  // not a closed manifold
        const uint32_t cur_edge = T[ct].e[ce];
        const uint32_t cur_vertex = T[ct].v[ce];
        T[ct].parent = new_face_id;
        if (!edges[cur_edge].deleted) {
          H.ledge.push_back(cur_edge);
          H.lvert.push_back(cur_vertex);
          num++;
          ce = (ce + 1) % 3;
        } else {
          // Edge::other_triangle, Triangle::next_edge_id
          const Edge& e = edges[cur_edge];
          ct = ct == e.f[0] ? e.f[1] : e.f[0];
          int nx = -1;
          for (int k = 0; k < 3; ++k)
            if (T[ct].e[k] == cur_edge) nx = (k + 1) % 3;
          if (nx < 0 || T[ct].v[nx] != cur_vertex) return false;
          ce = nx;
        }
      }
      if (num > 2) {
        H.ffirst.push_back(first);
        H.fnum.push_back(num);
        faces_tri.push_back(i);
      } else {
        // a face of two edges is not kept (its loop entries stay, unreferenced)
      }
      break;
    }
  }
  const size_t nfaces = H.ffirst.size();
  for (size_t f = 0; f < nfaces; ++f) {
    const F3 n = T[faces_tri[f]].n;
    H.fnx.push_back(n.x);
    H.fny.push_back(n.y);
    H.fnz.push_back(n.z);
  }
  for (auto& e : edges) {
    H.ev0.push_back(e.v[0]);
    H.ev1.push_back(e.v[1]);
    H.edx.push_back(e.dir.x);
    H.edy.push_back(e.dir.y);
    H.edz.push_back(e.dir.z);
  }
  // vertices: adjacent faces and the edge leaving each vertex in that face's loop
  H.vfirst.assign(npts, 0);
  H.vnum.assign(npts, 0);
  for (size_t f = 0; f < nfaces; ++f)
    for (uint32_t k = H.ffirst[f]; k < H.ffirst[f] + H.fnum[f]; ++k) H.vnum[H.lvert[k]]++;
  uint32_t total = 0;
  for (uint32_t v = 0; v < npts; ++v) {
    H.vfirst[v] = total;
    total += H.vnum[v];
    H.vnum[v] = 0;
  }
  H.vface.assign(total, 0);
  H.vedge.assign(total, 0);
  for (size_t f = 0; f < nfaces; ++f)
    for (uint32_t k = H.ffirst[f]; k < H.ffirst[f] + H.fnum[f]; ++k) {
      const uint32_t v = H.lvert[k];
      H.vface[H.vfirst[v] + H.vnum[v]] = (uint32_t)f;
      H.vedge[H.vfirst[v] + H.vnum[v]] = H.ledge[k];
      H.vnum[v]++;
    }
  int64_t num_valid_vertices = 0;
  for (uint32_t v = 0; v < npts; ++v) num_valid_vertices += H.vnum[v] != 0;
  // Euler's characteristic
  return num_valid_vertices + (int64_t)nfaces - num_valid_edges == 2;
}

extern "C" int ncx_convex_hull(const float* xyz, size_t npts, uint32_t* out_ids, size_t* n_out) {
  if (!xyz || !out_ids || !n_out || npts > (1u << 15)) return NCX_E_INVALID;
  try {
    HullBuilder hb;
    *n_out = (size_t)hb.build(xyz, (int)npts, out_ids);
  } catch (const std::bad_alloc&) {
    return NCX_E_CAPACITY;
  }
  return NCX_OK;
}

extern "C" int ncx_convex_hull_batch(const float* xyz, const uint32_t* offsets, const uint32_t* counts,
                                     size_t nclouds, uint32_t* out_ids, uint32_t* out_counts, int nthreads) {
  if (!xyz || !offsets || !counts || !out_ids || !out_counts) return NCX_E_INVALID;
  for (size_t i = 0; i < nclouds; ++i)
    if (counts[i] > (1u << 15)) return NCX_E_INVALID;
  if (nthreads < 1) nthreads = 1;
  std::vector<std::thread> th;
  std::vector<int> failed(nthreads, 0);
  size_t chunk = (nclouds + nthreads - 1) / nthreads;
  try {
    for (int t = 0; t < nthreads; ++t) {
      size_t b = t * chunk, e = std::min(nclouds, b + chunk);
      if (b >= e) break;
      th.emplace_back([=, &failed]() {
        try {
          HullBuilder hb;
          for (size_t i = b; i < e; ++i)
            out_counts[i] = (uint32_t)hb.build(xyz + 3 * (size_t)offsets[i], (int)counts[i], out_ids + offsets[i]);
        } catch (const std::bad_alloc&) {
          failed[t] = 1;
        }
      });
    }
  } catch (...) {
    for (auto& x : th) x.join();
    return NCX_E_CAPACITY;
  }
  for (auto& x : th) x.join();
  for (int f : failed)
    if (f) return NCX_E_CAPACITY;
  return NCX_OK;
}

extern "C" int ncx_hull_triangulate(const float* xyz, size_t npts, uint32_t* tris, size_t cap, size_t* ntri) {
  if (!xyz || !ntri || npts > (1u << 15)) return NCX_E_INVALID;
  try {
    std::vector<uint32_t> t;
    hull_triangulate(xyz, (uint32_t)npts, t);
    *ntri = t.size() / 3;
    if (!tris) return NCX_OK;
    if (t.size() > 3 * cap) return NCX_E_CAPACITY;
    std::memcpy(tris, t.data(), t.size() * sizeof(uint32_t));
  } catch (const std::bad_alloc&) {
    return NCX_E_CAPACITY;
  }
  return NCX_OK;
}

extern "C" int ncx_hull_polyhedron(const float* xyz, size_t npts, const uint32_t* tris, size_t ntri,
                                   float* face_normals, uint32_t* face_first, uint32_t* face_count, size_t* nfaces,
                                   uint32_t* loop_vertex, uint32_t* loop_edge, size_t* nloops, uint32_t* edge_vertices,
                                   float* edge_dirs, size_t* nedges, uint32_t* vertex_first, uint32_t* vertex_count,
                                   uint32_t* adj_face, uint32_t* adj_edge) {
  if (!xyz || !tris || !nfaces || !nloops || !nedges || npts > (1u << 15) || ntri > (1u << 17)) return NCX_E_INVALID;
  try {
    std::vector<float> X(npts), Y(npts), Z(npts);
    for (size_t i = 0; i < npts; ++i) {
      X[i] = xyz[3 * i];
      Y[i] = xyz[3 * i + 1];
      Z[i] = xyz[3 * i + 2];
    }
    HullPoly H;
    if (!hull_poly_build(X.data(), Y.data(), Z.data(), (uint32_t)npts, tris, (uint32_t)ntri, H)) {
      *nfaces = *nloops = *nedges = 0;
      return NCX_E_INVALID;
    }
    // capacities: faces <= ntri, loop and adjacency entries <= 3 ntri, edges <= 3 ntri
    *nfaces = H.ffirst.size();
    *nloops = H.lvert.size();
    *nedges = H.ev0.size();
    for (size_t f = 0; f < H.ffirst.size(); ++f) {
      if (face_normals) {
        face_normals[3 * f] = H.fnx[f];
        face_normals[3 * f + 1] = H.fny[f];
        face_normals[3 * f + 2] = H.fnz[f];
      }
      if (face_first) face_first[f] = H.ffirst[f];
      if (face_count) face_count[f] = H.fnum[f];
    }
    for (size_t k = 0; k < H.lvert.size(); ++k) {
      if (loop_vertex) loop_vertex[k] = H.lvert[k];
      if (loop_edge) loop_edge[k] = H.ledge[k];
    }
    for (size_t e = 0; e < H.ev0.size(); ++e) {
      if (edge_vertices) {
        edge_vertices[2 * e] = H.ev0[e];
        edge_vertices[2 * e + 1] = H.ev1[e];
      }
      if (edge_dirs) {
        edge_dirs[3 * e] = H.edx[e];
        edge_dirs[3 * e + 1] = H.edy[e];
        edge_dirs[3 * e + 2] = H.edz[e];
      }
    }
    for (size_t v = 0; v < npts; ++v) {
      if (vertex_first) vertex_first[v] = H.vfirst[v];
      if (vertex_count) vertex_count[v] = H.vnum[v];
    }
    for (size_t k = 0; k < H.vface.size(); ++k) {
      if (adj_face) adj_face[k] = H.vface[k];
      if (adj_edge) adj_edge[k] = H.vedge[k];
    }
  } catch (const std::bad_alloc&) {
    return NCX_E_CAPACITY;
  }
  return NCX_OK;
}

extern "C" int ncx_hull_polyhedra_check(const float* pool_xyz, const uint32_t* offsets, const uint32_t* counts,
                                        size_t nhulls, uint8_t* ok, int nthreads) {
  if (!pool_xyz || !offsets || !counts || !ok) return NCX_E_INVALID;
  for (size_t i = 0; i < nhulls; ++i)
    if (counts[i] > (1u << 15)) return NCX_E_INVALID;
  if (nthreads < 1) nthreads = 1;
  std::vector<std::thread> th;
  std::vector<int> failed(nthreads, 0);
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([=, &failed]() {
      try {
        std::vector<uint32_t> tris;
        std::vector<float> X, Y, Z;
        HullPoly H;
        for (size_t i = t; i < nhulls; i += nthreads) {
          const float* p = pool_xyz + 3 * (size_t)offsets[i];
          const uint32_t n = counts[i];
          ok[i] = 0;
          if (!hull_triangulate(p, n, tris)) continue;
          X.resize(n); Y.resize(n); Z.resize(n);
          for (uint32_t k = 0; k < n; ++k) { X[k] = p[3 * k]; Y[k] = p[3 * k + 1]; Z[k] = p[3 * k + 2]; }
          ok[i] = hull_poly_build(X.data(), Y.data(), Z.data(), n, tris.data(), (uint32_t)(tris.size() / 3), H) ? 1 : 0;
        }
      } catch (const std::bad_alloc&) {
        failed[t] = 1;
      }
    });
  for (auto& x : th) x.join();
  for (int f : failed)
    if (f) return NCX_E_CAPACITY;
  return NCX_OK;
}
