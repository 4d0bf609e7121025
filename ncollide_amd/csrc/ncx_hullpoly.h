// ncx_hullpoly.h -- the polyhedral structure of a convex hull (host code, input preparation).
//
// ConvexHull::try_new (src/shape/convex.rs:107-324) turns a closed triangle mesh into polygonal
// faces: an edge whose two triangles have normals closer than 1 - sqrt(f32::EPSILON) is deleted,
// and the faces are the contours walked around deleted edges.  The contact-manifold generators
// read the result: support_face_toward scans face normals, support_feature_toward walks the
// faces and edges adjacent to the support vertex (convex.rs:388-415, 487-545).
#pragma once
#include <cstdint>
#include <vector>

struct HullPoly {
  // faces: normal, and a run [first, first + num) of the loop arrays
  std::vector<float> fnx, fny, fnz;
  std::vector<uint32_t> ffirst, fnum;
  // loops (vertices_adj_to_face, edges_adj_to_face): the face's vertices in order, and the edge
  // from each to the next
  std::vector<uint32_t> lvert, ledge;
  // edges: end vertices (as first met) and unit direction v0 -> v1
  std::vector<uint32_t> ev0, ev1;
  std::vector<float> edx, edy, edz;
  // vertices: a run [first, first + num) of the adjacency arrays (faces_adj_to_vertex,
  // edges_adj_to_vertex)
  std::vector<uint32_t> vfirst, vnum, vface, vedge;
};

// Builds the structure of the hull whose vertices are pts[0..npts) (SoA) from its outward
// (counter-clockwise seen from outside) triangles.  Returns false where the reference's try_new
// returns None: a degenerate edge or triangle, or a result that breaks Euler's characteristic.
bool hull_poly_build(const float* x, const float* y, const float* z, uint32_t npts, const uint32_t* tris,
                     uint32_t ntri, HullPoly& out);

// The outward triangles of the convex hull of pts[0..npts) (interleaved xyz), indexing the input
// points; empty for a degenerate (flat) cloud.
bool hull_triangulate(const float* xyz, uint32_t npts, std::vector<uint32_t>& tris);
