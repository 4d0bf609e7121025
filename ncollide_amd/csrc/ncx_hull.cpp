// ncx_hull.cpp -- convex hull vertex selection (input preparation, host code).
//
// Stands in for ConvexHull::try_from_points -> transformation::convex_hull3
// (src/shape/convex.rs:88-97, src/transformation/convex_hull3.rs:28-149).  The
// reference runs quickhull on PCA-whitened points and maps the result back,
// which perturbs vertex coordinates by rounding and reorders them; here the
// hull is built incrementally in double precision and the result is the list of
// input indices that are hull vertices, in input order, so hull vertices are
// bit-identical to the input points.  Support functions only read the vertex
// list, so this is all the GPU queries need.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include "../../include/ncx.h"

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
  std::vector<int> edge_face;  // directed edge (a, b) -> face id, stamped
  std::vector<uint32_t> edge_stamp;
  uint32_t stamp = 0;
  int n = 0;
  std::vector<int> visible;
  std::vector<std::pair<int, int>> horizon;
  std::vector<uint8_t> is_vis;

  void reset_edges(int npts) {
    if ((int)edge_face.size() < npts * npts) {
      edge_face.assign((size_t)npts * npts, -1);
      edge_stamp.assign((size_t)npts * npts, 0);
      stamp = 0;
    }
    stamp++;
    if (stamp == 0) {
      std::fill(edge_stamp.begin(), edge_stamp.end(), 0);
      stamp = 1;
    }
  }
  int get_edge(int a, int b) const {
    size_t k = (size_t)a * n + b;
    return edge_stamp[k] == stamp ? edge_face[k] : -1;
  }
  void set_edge(int a, int b, int f) {
    size_t k = (size_t)a * n + b;
    edge_face[k] = f;
    edge_stamp[k] = stamp;
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

  // Returns the number of hull vertices written to `out` (input indices, ascending), 0 if degenerate.
  int build(const float* xyz, int npts, uint32_t* out) {
    if (npts < 4) return 0;
    n = npts;
    std::vector<D3> P(npts);
    double scale = 0.0;
    for (int i = 0; i < npts; ++i) {
      P[i] = D3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
      scale = std::max(scale, std::max(std::fabs(P[i].x), std::max(std::fabs(P[i].y), std::fabs(P[i].z))));
    }
    if (scale == 0.0) return 0;
    const double eps = 1e-11 * scale;
    // Initial tetrahedron: extreme point, farthest point, farthest from the line, farthest from the plane.
    int i0 = 0;
    for (int i = 1; i < npts; ++i) if (P[i].x < P[i0].x) i0 = i;
    int i1 = -1;
    double best = 0;
    for (int i = 0; i < npts; ++i) { double d = nrm(sub(P[i], P[i0])); if (d > best) { best = d; i1 = i; } }
    if (i1 < 0 || best <= eps) return 0;
    D3 l = sub(P[i1], P[i0]);
    int i2 = -1;
    best = 0;
    for (int i = 0; i < npts; ++i) { double d = nrm(crs(l, sub(P[i], P[i0]))); if (d > best) { best = d; i2 = i; } }
    if (i2 < 0 || best <= eps * nrm(l)) return 0;
    D3 pn = crs(l, sub(P[i2], P[i0]));
    int i3 = -1;
    best = 0;
    for (int i = 0; i < npts; ++i) { double d = std::fabs(dt(pn, sub(P[i], P[i0]))); if (d > best) { best = d; i3 = i; } }
    if (i3 < 0 || best <= eps * nrm(pn)) return 0;
    faces.clear();
    reset_edges(npts);
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
    std::vector<uint8_t> used(npts, 0);
    for (auto& f : faces)
      if (f.alive) used[f.v[0]] = used[f.v[1]] = used[f.v[2]] = 1;
    int k = 0;
    for (int i = 0; i < npts; ++i) if (used[i]) out[k++] = (uint32_t)i;
    return k;
  }
};

}  // namespace

extern "C" int ncx_convex_hull(const float* xyz, size_t npts, uint32_t* out_ids, size_t* n_out) {
  if (!xyz || !out_ids || !n_out || npts > (1u << 15)) return NCX_E_INVALID;
  HullBuilder hb;
  *n_out = (size_t)hb.build(xyz, (int)npts, out_ids);
  return NCX_OK;
}

extern "C" int ncx_convex_hull_batch(const float* xyz, const uint32_t* offsets, const uint32_t* counts,
                                     size_t nclouds, uint32_t* out_ids, uint32_t* out_counts, int nthreads) {
  if (!xyz || !offsets || !counts || !out_ids || !out_counts) return NCX_E_INVALID;
  for (size_t i = 0; i < nclouds; ++i)
    if (counts[i] > (1u << 15)) return NCX_E_INVALID;
  if (nthreads < 1) nthreads = 1;
  std::vector<std::thread> th;
  size_t chunk = (nclouds + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    size_t b = t * chunk, e = std::min(nclouds, b + chunk);
    if (b >= e) break;
    th.emplace_back([=]() {
      HullBuilder hb;
      for (size_t i = b; i < e; ++i)
        out_counts[i] = (uint32_t)hb.build(xyz + 3 * (size_t)offsets[i], (int)counts[i], out_ids + offsets[i]);
    });
  }
  for (auto& x : th) x.join();
  return NCX_OK;
}
