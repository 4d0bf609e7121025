/* ncx.h -- C ABI of the MI355X collision hot path (libncx.so).
 *
 * Batch-first replacement for ncollide3d 0.33's broad phase, convex contact /
 * distance queries and TriMesh ray casting.  Every function returns 0 on
 * success or a negative NCX_E_* code; ncx_last_error() gives a message.
 * Buffers named *_dev are device pointers (hipMalloc'd or torch CUDA tensors)
 * resident on the context's device; all work is enqueued on the context's
 * stream and is complete when the call returns unless noted.  Handles are
 * opaque; calls on one handle are not thread-safe (the reference takes
 * `&mut self`), distinct handles are independent.
 *
 * Layouts shared by several entry points:
 *   pose      : ncx_pose, 32 bytes: translation then unit quaternion (i, j, k, w),
 *               the fields of nalgebra's Isometry3<f32> (reference src/lib.rs:94-136).
 *   AABB SoA  : six float arrays minx, miny, minz, maxx, maxy, maxz (reference
 *               AABB{mins, maxs}, src/bounding_volume/aabb.rs:37-40).
 *   shape     : ncx_shape, 32 bytes; hull vertices live in a pool of three float
 *               arrays (x, y, z), local frame, indexed by hull_offset/hull_count.
 *   pair keys : (a << 32) | b with a < b, the reference's SortedPair
 *               (src/utils/sorted_pair.rs:7-26).
 */
#ifndef NCX_H
#define NCX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NCX_OK 0
#define NCX_E_INVALID (-1)    /* bad argument (null handle, size mismatch, bad shape type) */
#define NCX_E_HIP (-2)        /* HIP runtime error */
#define NCX_E_CAPACITY (-3)   /* caller-provided output capacity too small; *n holds the size needed */
#define NCX_E_NODEVICE (-4)   /* no GPU visible */

enum ncx_shape_type { NCX_BALL = 0, NCX_CUBOID = 1, NCX_CAPSULE = 2, NCX_CONVEX_HULL = 3 };

typedef struct ncx_pose {
  float tx, ty, tz, pad;
  float qi, qj, qk, qw;
} ncx_pose;

/* Ball: p0 = radius (src/shape/ball.rs); Cuboid: p0..p2 = half extents
 * (src/shape/cuboid.rs); Capsule: p0 = half height along local y, p1 = radius
 * (src/shape/capsule.rs); ConvexHull: hull_offset/hull_count into the pool
 * (src/shape/convex.rs). */
typedef struct ncx_shape {
  uint32_t type;
  uint32_t hull_offset;
  uint32_t hull_count;
  float p0, p1, p2;
  uint32_t pad[2];
} ncx_shape;

typedef struct ncx_ctx ncx_ctx;
typedef struct ncx_bp ncx_bp;
typedef struct ncx_shapes ncx_shapes;
typedef struct ncx_mesh ncx_mesh;

const char* ncx_last_error(void);

/* Context: device + stream.  `stream` is a hipStream_t (NULL = the default stream).
 * No reference counterpart: the reference is a single-threaded CPU library. */
int ncx_init(ncx_ctx** ctx, int device, void* stream);
int ncx_free(ncx_ctx* ctx);
int ncx_set_stream(ncx_ctx* ctx, void* stream);
int ncx_synchronize(ncx_ctx* ctx);

/* ---------------------------------------------------------------- broad phase
 * Replaces DBVTBroadPhase (src/pipeline/broad_phase/dbvt_broad_phase.rs:57-460)
 * behind the BroadPhase trait (src/pipeline/broad_phase/broad_phase.rs:41-86).
 * The stored box of each proxy follows the reference exactly: create_proxy
 * stores the box as given (:275-280); deferred_set_bounding_volume replaces it
 * by box.loosened(margin) only when the new box is not contained in the stored
 * one (:325-347).  After ncx_bp_update the pair set is
 * { (a,b) : a<b, stored_a intersects stored_b } (aabb.rs:156-158), the set the
 * reference's HashMap holds after update (:174-259); it is recomputed on the GPU
 * by a Morton-sorted LBVH instead of the incremental DBVT. */
int ncx_bp_create(ncx_ctx* ctx, float margin, ncx_bp** bp);                    /* DBVTBroadPhase::new (:81) */
int ncx_bp_destroy(ncx_bp* bp);
/* create_proxy for n boxes (host SoA); handles are slab keys (LIFO reuse, as the
 * slab crate behind `proxies: Slab<...>` at :58). */
int ncx_bp_create_proxies(ncx_bp* bp, const float* const aabb_soa_host[6], size_t n, uint32_t* handles_out);
/* deferred_set_bounding_volume for n proxies, boxes on the device (SoA). */
int ncx_bp_set_aabbs_dev(ncx_bp* bp, const uint32_t* handles_dev, const float* const aabb_soa_dev[6], size_t n);
/* Same with host arrays. */
int ncx_bp_set_aabbs(ncx_bp* bp, const uint32_t* handles, const float* const aabb_soa_host[6], size_t n);
/* remove (:282-323). */
int ncx_bp_remove(ncx_bp* bp, const uint32_t* handles, size_t n);
/* remove (:282-323) with its handler: the pairs of the current set that contain a removed proxy are
 * dropped and written to removed_keys (sorted, host; nullable), *n_removed = their number.  They
 * are not reported again as stopped by the next update's ncx_bp_events. */
int ncx_bp_remove_with_pairs(ncx_bp* bp, const uint32_t* handles, size_t n, uint64_t* removed_keys, size_t cap,
                             size_t* n_removed);
/* The pair filter BroadPhaseInterferenceHandler::is_interference_allowed (broad_phase.rs:31) in the
 * form the world gives it: per-proxy CollisionGroups (src/pipeline/object/collision_groups.rs:
 * membership, whitelist, blacklist masks over 30 groups); a pair is kept when
 * can_interact_with_groups (:353-359) holds.  Proxies start with CollisionGroups::new() (all
 * groups).  Host arrays; takes effect at the next update, which recomputes every pair. */
int ncx_bp_set_collision_groups(ncx_bp* bp, const uint32_t* handles, const uint32_t* membership,
                                const uint32_t* whitelist, const uint32_t* blacklist, size_t n);
/* interferences_with_bounding_volume / _with_point / _with_ray (:387-432), batched: the proxies of
 * the last update's tree (less those removed since) whose stored box intersects query box q
 * (aabb.rs:156-158), contains point q (aabb.rs:138-146), or is hit by ray q within max_toi
 * (AABB::intersects_ray, src/query/ray/ray_aabb.rs:13-50).  Queries are device SoA; the handles of
 * query q go to handles_dev[offsets_dev[q] .. offsets_dev[q + 1]) (offsets_dev: nq + 1 entries,
 * device), in no particular order.  *total = the number of handles; NCX_E_CAPACITY (nothing
 * written) when it exceeds cap. */
int ncx_bp_interferences_with_aabbs(ncx_bp* bp, const float* const aabb_soa_dev[6], size_t nq, uint32_t* offsets_dev,
                                    uint32_t* handles_dev, size_t cap, size_t* total);
int ncx_bp_interferences_with_points(ncx_bp* bp, const float* const point_soa_dev[3], size_t nq,
                                     uint32_t* offsets_dev, uint32_t* handles_dev, size_t cap, size_t* total);
int ncx_bp_interferences_with_rays(ncx_bp* bp, const float* const ray_soa_dev[6], size_t nq, float max_toi,
                                   uint32_t* offsets_dev, uint32_t* handles_dev, size_t cap, size_t* total);
/* update (:174-259): rebuild the pair set on the GPU. */
int ncx_bp_update(ncx_bp* bp);
/* num_interferences (:97-99). */
int ncx_bp_num_pairs(ncx_bp* bp, size_t* n);
/* Current pairs as sorted keys (host copy). Returns NCX_E_CAPACITY with *n set if cap is too small. */
int ncx_bp_pairs(ncx_bp* bp, uint64_t* keys_host, size_t cap, size_t* n);
/* Device pointers to the current pair list (a[i] < b[i]), valid until the next update. */
int ncx_bp_pairs_dev(ncx_bp* bp, const uint32_t** a_dev, const uint32_t** b_dev, size_t* n);
/* Copies the current (unsorted) pair list into caller device buffers of capacity cap (stream-ordered). */
int ncx_bp_copy_pairs_dev(ncx_bp* bp, uint32_t* a_dev, uint32_t* b_dev, size_t cap, size_t* n);
/* interference_started / interference_stopped of the last update (broad_phase.rs:28-38), sorted keys. */
int ncx_bp_events(ncx_bp* bp, uint64_t* started, size_t cap_started, size_t* n_started, uint64_t* stopped,
                  size_t cap_stopped, size_t* n_stopped);
/* The stored box of a proxy (BroadPhase::proxy, :262-273), 6 floats. Returns NCX_E_INVALID if absent. */
int ncx_bp_proxy_aabb(ncx_bp* bp, uint32_t handle, float out[6]);

/* ---------------------------------------------------------------- shapes
 * Uploads a shape set (host arrays) to the device.  Hull pools are local-frame
 * vertices (ConvexHull::points, src/shape/convex.rs:350-353). */
int ncx_shapes_create(ncx_ctx* ctx, const ncx_shape* shapes, size_t n, const float* hull_x, const float* hull_y,
                      const float* hull_z, size_t n_hull_points, ncx_shapes** out);
int ncx_shapes_destroy(ncx_shapes* s);
int ncx_shapes_count(ncx_shapes* s, size_t* n);
/* No reference counterpart: the autotuner's choice per vertex-count bucket (pairs by their larger
 * hull vertex count: none, <= 32, <= 96, more) for query 0 (contact) or 1 (distance): contact
 * 0, 1, 2 = 8, 16, 32 lanes per pair with the 44-vertex first-tier EPA arena, 3, 4 = 8, 16 lanes
 * with the 24-vertex arena; distance (no arena) 0, 1, 2, 3, 4 = 8, 16, 32, 4, 2 lanes per pair;
 * -1 before the first batch large enough to tune on;
 * sample_ms (nullable, 20 floats) the sample time of each variant per bucket (0 for a variant not
 * tried). */
int ncx_shapes_tuning(ncx_shapes* s, int query, int32_t* choice, float* sample_ms);
/* No reference counterpart: sets the first-tier variant per bucket (as ncx_shapes_tuning reports
 * them; -1 = tune on the next large batch).  Results do not depend on the variant. */
int ncx_shapes_set_tuning(ncx_shapes* s, int query, const int32_t* choice);

/* World AABB of every shape at the given poses (device arrays), optionally loosened by
 * `margin` (Shape::aabb: src/bounding_volume/aabb_ball.rs:8-13, aabb_cuboid.rs:7-14,
 * aabb_support_map.rs:35-45, aabb_convex.rs:7-11; AABB::loosened aabb.rs:190-199). */
int ncx_shapes_aabbs_dev(ncx_shapes* s, const ncx_pose* poses_dev, float margin, float* const aabb_soa_dev[6]);

/* ---------------------------------------------------------------- contact / distance
 * Batched query::contact (src/query/contact/contact_shape_shape.rs:10-48) over
 * pairs (ia[k], ib[k]) of one shape set at `poses_dev` (indexed by shape).
 * Dispatch, closed forms, GJK (src/query/algorithms/gjk.rs:76-177) and EPA
 * (src/query/algorithms/epa3.rs:219-430) follow the reference operation by
 * operation in f32.  Outputs (device): has_contact[k] = 1 for Some(Contact);
 * world1/world2/normal as 3 floats per pair; depth.  gjk_iters may be NULL.
 * init_dir_dev (nullable, 3 floats per pair): the `init_dir` argument of
 * contact_support_map_support_map_with_params
 * (src/query/contact/contact_support_map_support_map.rs:38-61) for pairs that
 * reach the support-map path; NULL is the reference's None, and so is an all-zero entry. */
int ncx_contact_batch(ncx_shapes* s, const ncx_pose* poses_dev, const uint32_t* ia_dev, const uint32_t* ib_dev,
                      size_t npairs, float prediction, const float* init_dir_dev, uint8_t* has_contact_dev,
                      float* world1_dev, float* world2_dev, float* normal_dev, float* depth_dev,
                      int32_t* gjk_iters_dev);
/* Batched query::distance (src/query/distance/distance.rs:9-33). */
int ncx_distance_batch(ncx_shapes* s, const ncx_pose* poses_dev, const uint32_t* ia_dev, const uint32_t* ib_dev,
                       size_t npairs, float* dist_dev, int32_t* gjk_iters_dev);

/* ---------------------------------------------------------------- time of impact
 * Batched query::time_of_impact (src/query/time_of_impact/time_of_impact.rs:66-158) under
 * translational motion over pairs (ia[k], ib[k]) of one shape set: velocities per shape (vel_dev,
 * 3 floats per shape, like poses_dev).  Ball-Ball: time_of_impact_ball_ball
 * (time_of_impact_ball_ball.rs:7-61); every other pair: time_of_impact_support_map_support_map
 * (time_of_impact_support_map_support_map.rs:10-116), GJK's ray cast on the Minkowski difference
 * (src/query/algorithms/gjk.rs:214-365).  Outputs per pair: has = 1 for Some(TOI); toi; witness1,
 * witness2, normal1, normal2 (3 floats each, local frames of the two shapes); status 0
 * OutOfIterations, 1 Converged, 2 Failed, 3 Penetrating (time_of_impact.rs:7-25). */
int ncx_time_of_impact_batch(ncx_shapes* s, const ncx_pose* poses_dev, const float* vel_dev, const uint32_t* ia_dev,
                             const uint32_t* ib_dev, size_t npairs, float max_toi, float target_distance,
                             uint8_t* has_dev, float* toi_dev, float* witness1_dev, float* witness2_dev,
                             float* normal1_dev, float* normal2_dev, uint8_t* status_dev);
/* Batched RayCast::toi_and_normal_with_ray (src/query/ray/ray.rs:125-160) of shape shape_dev[k] at
 * its pose against ray k (device SoA origins and directions): Ball (ray_ball.rs:32-43), Cuboid
 * (ray_cuboid.rs:17-26 -> ray_aabb.rs:45-61, 178-290), Capsule and ConvexHull (ray_support_map.rs:
 * 15-62, 117-163, GJK ray cast).  Hit: toi, world normal, feature = face id (cuboid faces 0-5, 0 for
 * a ball, 0xFFFFFFFE for FeatureId::Unknown); miss: toi = -1, feature = UINT32_MAX. */
int ncx_shapes_cast_rays(ncx_shapes* s, const ncx_pose* poses_dev, const uint32_t* shape_dev,
                         const float* const ray_soa_dev[6], size_t n, float max_toi, int solid, float* toi_dev,
                         float* normal_dev, uint32_t* feature_dev);

/* ---------------------------------------------------------------- TriMesh ray casting
 * TriMesh::new (src/shape/trimesh.rs:100-197) with a binned-SAH tree (built on the host, then
 * a 4-wide cast copy on the device) in place of the median-split BVT; host arrays. */
int ncx_trimesh_create(ncx_ctx* ctx, const float* points_xyz, size_t npoints, const uint32_t* indices,
                       size_t ntriangles, ncx_mesh** mesh);
int ncx_trimesh_destroy(ncx_mesh* mesh);
/* RayCast::toi_and_normal_with_ray for TriMesh (src/query/ray/ray_trimesh.rs:23-50) on
 * n rays (device SoA origins ox/oy/oz and directions dx/dy/dz), mesh at `pose`.
 * Hit: toi, world normal (3 floats per ray), feature = face id, or face id +
 * ntriangles for a back-face hit (ray_trimesh.rs:41-47).  Miss: feature =
 * UINT32_MAX, toi = -1.  The hit kept is the least (toi, feature) over all triangles; the
 * reference's best-first search (src/partitioning/bvh.rs:101-160) can stop on a hit 1-2 ulps
 * farther when a ray passes exactly through a mesh vertex. */
int ncx_trimesh_cast_rays(ncx_mesh* mesh, const ncx_pose* pose, const float* const ray_soa_dev[6], size_t n,
                          float max_toi, float* toi_dev, float* normal_dev, uint32_t* feature_dev);
/* Same as ncx_trimesh_cast_rays, but only enqueues the kernel on the context stream (no host
 * synchronisation); ncx_trimesh_check reports a traversal failure of any earlier cast. */
int ncx_trimesh_cast_rays_async(ncx_mesh* mesh, const ncx_pose* pose, const float* const ray_soa_dev[6], size_t n,
                                float max_toi, float* toi_dev, float* normal_dev, uint32_t* feature_dev);
int ncx_trimesh_check(ncx_mesh* mesh);
/* No reference counterpart: the same cast as ncx_trimesh_cast_rays_async for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no allocating routine: its work counters and traversal spill stacks live
 * in a 16 MiB device array of the library, one per GPU, held until the process ends.  Enqueue
 * only; ncx_trimesh_check reports a traversal failure.  Every call on one GPU shares that array,
 * so calls that overlap on one GPU (different streams or host threads) race.  Returns
 * NCX_E_CAPACITY when the mesh's tree is too deep for the scratch to hold one block of spill
 * stacks per CU. */
int ncx_trimesh_cast_rays_noalloc(ncx_mesh* mesh, const ncx_pose* pose, const float* const ray_soa_dev[6], size_t n,
                                  float max_toi, float* toi_dev, float* normal_dev, uint32_t* feature_dev);
/* No reference counterpart (measurement): record events around each cast's traversal kernel;
 * ncx_trimesh_kernel_ms returns the last cast's kernel time once it has completed. */
int ncx_trimesh_set_timing(ncx_mesh* mesh, int enable);
int ncx_trimesh_kernel_ms(ncx_mesh* mesh, float* ms);
/* Traversal statistics of the same cast (nodes visited, triangles tested per ray), for roofline accounting. */
int ncx_trimesh_cast_stats(ncx_mesh* mesh, const ncx_pose* pose, const float* const ray_soa_dev[6], size_t n,
                           float max_toi, uint32_t* nodes_visited_dev, uint32_t* tris_tested_dev);

/* ---------------------------------------------------------------- composite shape (TriMesh) against convex shapes
 * contact_composite_shape_shape (src/query/contact/contact_composite_shape_shape.rs:9-45) of the mesh
 * at `mesh_pose` (host) and shape shape_dev[k] of `s` at poses_dev: the triangles whose box
 * intersects the shape's box in the mesh frame loosened by `prediction`, query::contact(mesh pose,
 * triangle, shape pose, shape, prediction) on each (a ball takes the closed-form point projection on
 * the triangle, contact_ball_convex_polyhedron.rs:11-77; every other shape GJK/EPA with the triangle's
 * support map, triangle.rs:290-307), and the deepest contact kept (world points, normal from the
 * mesh toward the shape, depth; face_dev (nullable): its triangle id, UINT32_MAX when none).  The
 * reference keeps the first of equally deep contacts in its BVT's order; here the least triangle id.
 * The mesh and the shape set must live on the same device. */
int ncx_mesh_contact_batch(ncx_mesh* mesh, const ncx_pose* mesh_pose, ncx_shapes* s, const ncx_pose* poses_dev,
                           const uint32_t* shape_dev, size_t n, float prediction, uint8_t* has_dev,
                           float* world1_dev, float* world2_dev, float* normal_dev, float* depth_dev,
                           uint32_t* face_dev);
/* distance_composite_shape_shape (src/query/distance/distance_composite_shape_shape.rs:9-101): the
 * least query::distance (GJK) between the mesh and shape shape_dev[k], found by a search of the mesh's
 * tree that skips boxes whose Minkowski-sum distance is not below the best so far. */
int ncx_mesh_distance_batch(ncx_mesh* mesh, const ncx_pose* mesh_pose, ncx_shapes* s, const ncx_pose* poses_dev,
                            const uint32_t* shape_dev, size_t n, float* dist_dev);

/* ---------------------------------------------------------------- input preparation (host)
 * ConvexHull::try_from_points (src/shape/convex.rs:88-97 -> transformation::convex_hull3,
 * src/transformation/convex_hull3.rs:28-149): writes the indices of the input points that are
 * vertices of their convex hull, ascending; *n_out = 0 for a degenerate (flat) cloud. */
int ncx_convex_hull(const float* points_xyz, size_t npoints, uint32_t* out_ids, size_t* n_out);
/* Many clouds on host threads: cloud i is points[offsets[i] .. offsets[i] + counts[i]); its hull
 * vertex ids (local to the cloud) go to out_ids[offsets[i] ..], their number to out_counts[i]. */
int ncx_convex_hull_batch(const float* points_xyz, const uint32_t* offsets, const uint32_t* counts, size_t nclouds,
                          uint32_t* out_ids, uint32_t* out_counts, int nthreads);
/* The outward triangles (counter-clockwise seen from outside) of the convex hull of the points,
 * indexing the input: the triangulation the contact-manifold generators' hull faces are built
 * from (ConvexHull::try_from_points hands convex_hull3's triangles to try_new,
 * src/shape/convex.rs:88-97).  tris = NULL: *ntri only.  *ntri = 0 for a flat cloud. */
int ncx_hull_triangulate(const float* points_xyz, size_t npoints, uint32_t* tris, size_t cap, size_t* ntri);
/* Whether ConvexHull::try_new succeeds on each hull of a pool (hull i: its vertices
 * pool[offsets[i] .. offsets[i] + counts[i]), triangulated as ncx_hull_triangulate): ok[i] = 1
 * or 0.  The reference has no ConvexHull where it fails (try_from_points returns None); the
 * contact-manifold generators need its faces. */
int ncx_hull_polyhedra_check(const float* pool_xyz, const uint32_t* offsets, const uint32_t* counts, size_t nhulls,
                             uint8_t* ok, int nthreads);
/* ConvexHull::try_new (src/shape/convex.rs:107-324) on a hull's triangles: polygonal faces
 * (normal, run of the loop arrays), loops (vertex, outgoing edge), edges (end vertices, unit
 * direction), per vertex a run of the adjacency arrays (face, the vertex's outgoing edge in that
 * face).  Capacities: faces <= ntri, loop and adjacency entries <= 3 ntri, edges <= 3 ntri.
 * NCX_E_INVALID where try_new returns None.  Any output pointer may be NULL. */
int ncx_hull_polyhedron(const float* points_xyz, size_t npoints, const uint32_t* tris, size_t ntri,
                        float* face_normals, uint32_t* face_first, uint32_t* face_count, size_t* nfaces,
                        uint32_t* loop_vertex, uint32_t* loop_edge, size_t* nloops, uint32_t* edge_vertices,
                        float* edge_dirs, size_t* nedges, uint32_t* vertex_first, uint32_t* vertex_count,
                        uint32_t* adj_face, uint32_t* adj_edge);

/* Warm-started contact batch: like ncx_contact_batch with init_dir, but GJK from init_dir stops
 * after max_iter iterations and such a pair is recomputed with init_dir = None (plain GJK).
 * No reference counterpart beyond the init_dir hook of
 * src/query/contact/contact_support_map_support_map.rs:38-61. */
int ncx_contact_batch_capped(ncx_shapes* s, const ncx_pose* poses_dev, const uint32_t* ia_dev,
                             const uint32_t* ib_dev, size_t npairs, float prediction, const float* init_dir_dev,
                             int32_t max_iter, uint8_t* has_contact_dev, float* world1_dev, float* world2_dev,
                             float* normal_dev, float* depth_dev, int32_t* gjk_iters_dev);
/* Warm-started distance batch: the init_dir argument of
 * distance_support_map_support_map_with_params (src/query/distance/distance_support_map_support_map.rs:25-53)
 * per pair (NULL or an all-zero entry = None), with GJK from init_dir stopped after max_iter
 * iterations and such a pair recomputed from the default direction (plain GJK). */
int ncx_distance_batch_capped(ncx_shapes* s, const ncx_pose* poses_dev, const uint32_t* ia_dev,
                              const uint32_t* ib_dev, size_t npairs, const float* init_dir_dev, int32_t max_iter,
                              float* dist_dev, int32_t* gjk_iters_dev);

/* ---------------------------------------------------------------- learned initial direction
 * No reference counterpart: a 44-64-64-3 MLP (FP16 operands, FP32 accumulation) that predicts
 * the init_dir argument of contact_support_map_support_map_with_params
 * (src/query/contact/contact_support_map_support_map.rs:38-61).  Features of pair (a, b):
 * R_a^T R_b (9), R_a^T (t_b - t_a) (3), 16 local support values of a and of b
 * (ncx_shape_descriptors), padded to 48 FP16 per pair.  Parameters: a flat FP32 vector of
 * 8336 entries (W1 64x48, b1 64, W2 64x64, b2 64, W3 16x64, b3 16; padding stays zero). */
typedef struct ncx_mlp ncx_mlp;
int ncx_mlp_create(ncx_ctx* ctx, int in, int hidden, int out, uint64_t seed, ncx_mlp** mlp);
int ncx_mlp_destroy(ncx_mlp* mlp);
int ncx_mlp_num_params(ncx_mlp* mlp, size_t* n);
int ncx_mlp_get_params(ncx_mlp* mlp, float* params_host);
int ncx_mlp_set_params(ncx_mlp* mlp, const float* params_host);
/* 16 support values per shape along fixed Fibonacci directions of its local frame (16 floats per shape). */
int ncx_shape_descriptors(ncx_shapes* s, float* desc_dev);
/* Features (48 FP16 per pair) of pairs (ia[k], ib[k]). */
int ncx_features(ncx_shapes* s, const ncx_pose* poses_dev, const float* desc_dev, const uint32_t* ia_dev,
                 const uint32_t* ib_dev, size_t npairs, void* feat_f16_dev);
/* Raw network output (3 floats per row, shape-a frame). */
int ncx_mlp_forward(ncx_mlp* mlp, const void* feat_f16_dev, size_t n, float* out_dev);
/* Mean squared error (1/n) sum |f(x) - y|^2: its gradient (8336 floats) and value, on the device. */
int ncx_mlp_grad(ncx_mlp* mlp, const void* feat_f16_dev, const float* target_dev, size_t n, float* grad_dev,
                 float* loss_dev);
/* One Adam step (lr; beta 0.9 / 0.999) on the batch; *loss_host (nullable) gets the batch loss. */
int ncx_mlp_train_step(ncx_mlp* mlp, const void* feat_f16_dev, const float* target_dev, size_t n, float lr,
                       float* loss_host);
/* Features + forward + rotation to world + normalisation, fused: unit init_dir per pair. */
int ncx_mlp_predict_dirs(ncx_mlp* mlp, const ncx_pose* poses_dev, const float* desc_dev, const uint32_t* ia_dev,
                         const uint32_t* ib_dev, size_t npairs, float* dirs_dev);

/* ---------------------------------------------------------------- one collision step of a scene
 * The hot path of CollisionWorld::update (src/pipeline/world.rs:104-119) for a shape set whose
 * shapes are all broad-phase proxies: world AABBs loosened by `prediction`, the DBVTBroadPhase
 * stored-box rule with `margin` (dbvt_broad_phase.rs:325-347; the first step stores boxes as
 * given, :275-280), LBVH refit (rebuilt every k steps), every intersecting stored-box pair, and
 * query::contact at `prediction` on each pair.  No host synchronisation inside a step: with
 * use_graph != 0, refit steps and rebuild steps (after the first step) each replay a captured
 * hipGraph.  Outputs are capacity-`max_pairs` device arrays.  Device memory: about 2.3 KB per
 * shape (2 KB of it buffers each proxy's first 512 pair partners for the pair search) plus the
 * per-pair outputs and contact work lists. */
typedef struct ncx_world ncx_world;
int ncx_world_create(ncx_ctx* ctx, ncx_shapes* shapes, float margin, float prediction, size_t max_pairs,
                     ncx_world** world);
int ncx_world_destroy(ncx_world* world);
/* Learned init_dir (ncx_mlp) for polytope-polytope pairs, GJK capped at 32 iterations with plain
 * GJK on the cap; NULL turns it off. */
int ncx_world_set_warmstart(ncx_world* world, ncx_mlp* mlp);
/* The init_dir of each pair of the last step under the learned warm start (3 floats per pair, in
 * the order of ncx_world_outputs; all-zero = None), valid until the next step. */
int ncx_world_warmstart_dirs(ncx_world* world, const float** dirs_dev);
/* The persistent narrow phase's warm start: ConvexPolyhedronConvexPolyhedronManifoldGenerator's
 * last_gjk_dir (src/pipeline/narrow_phase/contact_generator/
 * convex_polyhedron_convex_polyhedron_manifold_generator.rs:24-33, 96-145) kept per pair of
 * polyhedra (Cuboid / ConvexHull on both sides) across steps on the device: a pair's GJK starts from
 * its previous step's result direction (GJK uncapped, as the reference), a pair new to the broad
 * phase starts from None.  Exclusive with ncx_world_set_warmstart.  Device memory: 24 bytes per
 * table slot (a power of two, at least 2 * max_pairs) plus 24 bytes per pair. */
int ncx_world_set_temporal_warmstart(ncx_world* world, int enable);
/* Under the temporal warm start: each pair's init_dir of the last step (all-zero = None) and its
 * result direction (the next step's last_gjk_dir), 3 floats per pair in the order of
 * ncx_world_outputs; valid after ncx_world_sync until the next step. */
int ncx_world_temporal_dirs(ncx_world* world, const float** init_dev, const float** result_dev);
/* LBVH rebuilt every k >= 1 steps and refit in between (default 2). */
int ncx_world_set_rebuild_every(ncx_world* world, int k);
int ncx_world_step(ncx_world* world, const ncx_pose* poses_dev, int use_graph);
/* No reference counterpart: turns each step's narrow phase on (default) or off, so the broad
 * phase can be run and timed alone. */
int ncx_world_set_contacts(ncx_world* world, int enable);
/* No reference counterpart (measurement): internal nodes visited by one pair-count traversal of
 * the current tree, the per-pass work of the step's pair search.  Synchronous. */
int ncx_world_query_stats(ncx_world* world, unsigned long long* nodes_visited);
/* No reference counterpart (measurement): of the last step's pair search, out[0] = leaves with
 * more than 512 later partners, out[1] = of those, the leaves whose partners past the 512th went
 * to the spill chunks (the others were searched a second time).  Waits for the world's stream. */
int ncx_world_spill_stats(ncx_world* world, uint32_t out[2]);
/* Waits for the last step: *npairs = pair count; NCX_E_CAPACITY if it exceeded max_pairs.
 * Every contact is complete on return: pairs whose EPA polytope outgrew the step's LDS arena
 * tiers are continued here with device-memory arenas, so the step's poses must stay unchanged
 * until this returns.  *n_epa_overflow = pairs of the step that needed more than the first tier. */
int ncx_world_sync(ncx_world* world, size_t* npairs, size_t* n_epa_overflow);
int ncx_world_outputs(ncx_world* world, uint32_t** ia, uint32_t** ib, uint8_t** has_contact, float** world1,
                      float** world2, float** normal, float** depth, int32_t** gjk_iters);
int ncx_world_stored_aabbs(ncx_world* world, const float** soa_dev);
/* World-level queries after a step, through the scene's broad phase (handles are shape indices).
 * ncx_world_interferences_with_rays: glue::interferences_with_ray (src/pipeline/glue/query.rs:
 * 13-66): the shapes whose stored box the ray enters within max_toi and whose shape (at poses_dev,
 * the last step's poses) the ray hits, RayCast::toi_and_normal_with_ray with solid = true; per hit
 * the handle, toi and world normal (toi_dev / normal_dev nullable), in CSR form as the
 * ncx_bp_interferences_* calls.  aabb_only != 0 gives BroadPhase::interferences_with_ray (stored
 * boxes only; poses_dev may be NULL).
 * ncx_world_first_interference_with_rays: first_interference_with_ray (glue/query.rs:200-224 ->
 * dbvt_broad_phase.rs:433-460, RayIntersectionCostFnVisitor ray_intersection_cost_fn_visitor.rs:
 * 64-107): per ray the hit with the least toi (then the least handle: the reference keeps the
 * first of equal-toi hits its heap pops), handle UINT32_MAX and toi -1 on a miss.  Batches of
 * 1,024 rays or more are searched in an order sorted by origin and direction octant, with
 * about 16 bytes per ray of sort buffers kept by the world; results go to each ray's own slot.
 * ncx_world_interferences_with_aabbs: glue::interferences_with_aabb (glue/query.rs) = the stored
 * boxes intersecting each query box. */
int ncx_world_interferences_with_rays(ncx_world* world, const ncx_pose* poses_dev, const float* const ray_soa_dev[6],
                                      size_t n, float max_toi, int aabb_only, uint32_t* offsets_dev,
                                      uint32_t* handles_dev, float* toi_dev, float* normal_dev, size_t cap,
                                      size_t* total);
int ncx_world_first_interference_with_rays(ncx_world* world, const ncx_pose* poses_dev,
                                           const float* const ray_soa_dev[6], size_t n, float max_toi,
                                           uint32_t* handle_dev, float* toi_dev, float* normal_dev);
int ncx_world_interferences_with_aabbs(ncx_world* world, const float* const aabb_soa_dev[6], size_t n,
                                       uint32_t* offsets_dev, uint32_t* handles_dev, size_t cap, size_t* total);

/* ---------------------------------------------------------------- contact manifolds
 * The narrow phase's persistent contact-manifold generators for Ball, Cuboid, Capsule and
 * ConvexHull pairs (src/pipeline/narrow_phase/contact_generator/: DefaultContactDispatcher,
 * default_contact_dispatcher.rs:27-97; BallBall-, BallConvexPolyhedron-, CapsuleShape-,
 * CapsuleCapsule- and ConvexPolyhedronConvexPolyhedronManifoldGenerator) and the ContactManifold
 * they fill (src/query/contact/contact_manifold.rs, DistanceBased(0.02) tracking), one update
 * per pair per call as NarrowPhase::update_contact (src/pipeline/narrow_phase/narrow_phase.rs:56-80).
 *
 * State: one slot per interaction (a generator and its manifold in the reference): the
 * polyhedron generator's last_gjk_dir and the tracking points of the last update's contacts.
 * The caller maps pairs to slots (e.g. by its pair table) and resets a slot when its interaction
 * starts.  Slots must be distinct within one call.
 *
 * Feature ids (FeatureId): kind in the top two bits (0 vertex, 1 edge, 2 face, 3 unknown), the
 * shape-local id below; a capsule's features are its preprocessor's (Face 0/1 the end caps,
 * Face 2 the side).  Each pair's contacts are written in the manifold's cache order at
 * [k * NCX_MANIFOLD_MAX, k * NCX_MANIFOLD_MAX + count[k]) of world1/world2/normal (xyz each),
 * depth, feature1, feature2 (device arrays); the reference iterates the same set in slab order.
 * A hull face of more than 16 vertices, or a manifold of more than NCX_MANIFOLD_MAX tracked
 * contacts, fails the call with NCX_E_CAPACITY. */
#define NCX_MANIFOLD_MAX 32
typedef struct ncx_manifolds ncx_manifolds;
int ncx_manifolds_create(ncx_ctx* ctx, size_t nslots, ncx_manifolds** out);
int ncx_manifolds_destroy(ncx_manifolds* m);
/* New interactions: clears last_gjk_dir and the tracked contacts of the slots (device array). */
int ncx_manifolds_reset(ncx_manifolds* m, const uint32_t* slots_dev, size_t n);
/* last_gjk_dir of the slots (device arrays; zero rows: None). */
int ncx_manifolds_dirs(ncx_manifolds* m, const uint32_t* slots_dev, size_t n, float* dirs_dev);
/* One ContactManifoldGenerator::generate_contacts per pair, prediction (linear, angular1,
 * angular2) as ContactPrediction. */
int ncx_contact_manifolds(ncx_manifolds* m, ncx_shapes* s, const ncx_pose* poses_dev, const uint32_t* ia_dev,
                          const uint32_t* ib_dev, const uint32_t* slots_dev, size_t npairs, float linear,
                          float angular1, float angular2, uint32_t* count_dev, float* world1_dev, float* world2_dev,
                          float* normal_dev, float* depth_dev, uint32_t* feature1_dev, uint32_t* feature2_dev);

#ifdef __cplusplus
}
#endif
#endif /* NCX_H */
