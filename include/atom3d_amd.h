/*
 * atom3d_amd — C ABI of the MI355X (gfx950) geometry library.
 *
 * Drop-in boundary for the hot path of Atom3D (Luo-Yihao/Atom3d).  The
 * reference binds three JIT-built pybind11 modules (bvh_cuda, cumtv_cuda,
 * floodfill_cuda); each entry point below replaces one of their functions,
 * cited as file:line into the reference.  The Python package atom3d_amd binds
 * these symbols with ctypes.
 *
 * Conventions (all entry points):
 *   - return 0 on success, a negative A3D_ERR_* code on failure; the message
 *     is kept per thread and read with a3d_last_error();
 *   - every pointer argument is a DEVICE pointer unless stated otherwise;
 *     sizes are int64_t element counts;
 *   - work is enqueued asynchronously on the given HIP stream (the caller's
 *     current stream); nothing is allocated by the library: outputs, the BVH
 *     blob and workspaces are buffers the caller owns and sizes with the
 *     *_bytes queries;
 *   - no function synchronises the device except where documented.
 */
#ifndef ATOM3D_AMD_H
#define ATOM3D_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* a3d_stream_t; /* == hipStream_t */

enum {
  A3D_OK = 0,
  A3D_ERR_ARG = -1,      /* invalid argument (null pointer, bad size, bad mode) */
  A3D_ERR_HIP = -2,      /* a HIP runtime call failed */
  A3D_ERR_CAPACITY = -3, /* caller buffer too small */
  A3D_ERR_FORMAT = -4,   /* blob is not a BVH built by this library */
  A3D_ERR_INTERNAL = -5  /* an internal bound was exceeded (message says which) */
};

/* Last error message of the calling thread ("" if none). */
const char* a3d_last_error(void);
/* Library version string, e.g. "atom3d_amd 0.1.0 gfx950". */
const char* a3d_version(void);

/* ------------------------------------------------------------------ BVH build
 * Replaces build_bvh (bvh_kernels.cu:860-925, bound at :1245) and
 * BVHAccelerator.__init__ (kernels/bvh.py:66-93).  The reference builds a
 * binned-SAH tree on the host; this builds a Morton LBVH on the device
 * (63-bit Morton keys of face centroids, Karras 2012) with leaves of up to
 * leaf_max <= 8 triangles, then (treelet_passes > 0) restructures the
 * internal nodes above the leaves for SAH cost, 7-leaf treelets (Karras &
 * Aila 2013).  The query results of every entry point below do not depend on
 * the tree; only their speed does.
 *
 * a3d_bvh_sizes: bytes of the BVH blob and of the build workspace for nF faces.
 * a3d_bvh_build_ex: V float32[nV*3], F int32[nF*3] (device) -> blob.
 *   treelet_passes in [0, 8].  Synchronises the stream once at the end to
 *   check the tree depth against the walks' stacks (A3D_ERR_CAPACITY beyond
 *   99 internal levels; a Morton tree stays within 96), and with > 0 passes
 *   a few times per pass (the number of rounds a pass needs is read back).
 *   A refined tree deeper than 96 internal levels is discarded for the
 *   plain one.
 * a3d_bvh_build: a3d_bvh_build_ex with A3D_BVH_TREELET_PASSES passes.
 */
#define A3D_BVH_TREELET_PASSES 0
int a3d_bvh_sizes(int64_t nF, int leaf_max, int64_t* blob_bytes, int64_t* ws_bytes);
int a3d_bvh_build_ex(const float* V, int64_t nV, const int32_t* F, int64_t nF, int leaf_max,
                     int treelet_passes, void* blob, int64_t blob_bytes, void* ws, int64_t ws_bytes,
                     a3d_stream_t stream);
int a3d_bvh_build(const float* V, int64_t nV, const int32_t* F, int64_t nF, int leaf_max,
                  void* blob, int64_t blob_bytes, void* ws, int64_t ws_bytes, a3d_stream_t stream);
/* Copies the blob header fields to HOST memory (synchronises the stream):
 * out[0]=n_tris, out[1]=n_nodes, out[2]=depth (internal nodes on the longest
 * root-to-leaf path), out[3]=leaf_max. */
int a3d_bvh_info(const void* blob, int64_t* out4, a3d_stream_t stream);

/* ------------------------------------------------------------------ UDF
 * Replaces bvh_udf (bvh_kernels.cu:1028-1066 / bvh_udf_kernel :363-463, bound
 * at :1246) and BVHAccelerator.udf (kernels/bvh.py:95-117).
 * P float32[K*3] -> dist float32[K], face int32[K] (original face index, -1
 * if none), closest float32[K*3], uvw float32[K*3] (nullable; zeros when
 * face == -1).  Closest face = minimum of the reference's FP32 squared
 * distance over ALL faces, ties to the lowest face index; results do not
 * depend on perm, the flags or the tree.
 * perm (nullable, uint32[K]): process the queries in this order (a
 * permutation of 0..K-1, e.g. from a3d_query_order); results are still
 * written at each query's own index.  flags: A3D_UDF_LANE = one independent
 * walk per lane instead of one shared walk per wavefront; A3D_UDF_NO_OBBS =
 * skip the children's oriented-box bounds (kept for comparisons).  counters (nullable,
 * device int64[5], accumulated, wave walk only): wavefronts, node visits,
 * leaf triangles seen, exact triangle tests run (all summed over wavefronts)
 * and triangles whose bound admitted them per lane (summed over lanes).
 *
 * a3d_query_order: perm = the queries sorted along a Morton curve of their
 * bounding box (the top 24 bits of a 30-bit code; stable: ties keep their
 * input order), so each wavefront covers a compact region.
 * Workspace of a3d_query_order_ws_bytes(K) bytes, 256-byte aligned; K < 2^31.
 */
#define A3D_UDF_LANE 1
#define A3D_UDF_DIRECT 2 /* wave walk testing each needed triangle at once (no exact-test queue) */
#define A3D_UDF_NO_OBBS 4 /* prune by the axis-aligned boxes only, ignoring the blob's oriented boxes (same results) */
#define A3D_UDF_SPLIT 8 /* few queries: 4 lanes per query, each walking a quarter of the tree (same results; no seed) */
#define A3D_UDF_REDO 16 /* a3d_bvh_udf_seeded: walk the queries left at face -1 again without a seed */
int a3d_query_order_ws_bytes(int64_t K, int64_t* bytes);
int a3d_query_order(const float* P, int64_t K, uint32_t* perm, void* ws, int64_t ws_bytes,
                    a3d_stream_t stream);
/* a3d_query_order_work: a3d_query_order's Morton order with its 64-query
 * packets (perm[64 j .. 64 j + 63], j < K / 64) regrouped by estimated cost,
 * heaviest first, the Morton order kept within each cost class; the
 * trailing partial packet stays last.  A packet's class is floor(log2(m +
 * 1)), m the largest work[] over the 8 grid points around each of its
 * queries (the corners of the query's cell of the grid lo3 + (i, j, k) h, i,
 * j, k in [0, R], clamped into it as a3d_udf_grid_seed does); work uint32
 * per grid point, a3d_bvh_udf_work's output for a3d_grid_points's points.
 * A packet far costlier than the rest would end the walk late if started
 * late: the cost order starts those first.  Workspace of
 * a3d_query_order_work_ws_bytes(K) bytes, 256-byte aligned; K < 2^31.  No
 * counterpart in the reference. */
int a3d_query_order_work_ws_bytes(int64_t K, int64_t* bytes);
int a3d_query_order_work(const float* P, int64_t K, const uint32_t* work, int R, const float* lo3, float h,
                         uint32_t* perm, void* ws, int64_t ws_bytes, a3d_stream_t stream);
/* a3d_query_order_noalloc: a3d_query_order (same perm, bit for bit) that
 * allocates nothing and takes no workspace: the workspace is a file-scope
 * device array of 48 MiB that each GPU holds until the process ends.  For
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc); with a3d_bvh_udf it gives the
 * ordered closest-point path without allocation.  K <= 2^21 (else
 * A3D_ERR_CAPACITY).  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (concurrent streams or threads) race.  No
 * counterpart in the reference. */
int a3d_query_order_noalloc(const float* P, int64_t K, uint32_t* perm, a3d_stream_t stream);
int a3d_bvh_udf(const void* blob, const float* P, int64_t K, const uint32_t* perm, float* dist,
                int32_t* face, float* closest, float* uvw, int flags, int64_t* counters,
                a3d_stream_t stream);
/* a3d_bvh_udf_work: a3d_bvh_udf (wave walk; flags 0, A3D_UDF_DIRECT or
 * A3D_UDF_NO_OBBS) that also writes work[q] (uint32): the number of leaf
 * triangles whose bound admitted query q during its packet's walk, the
 * per-query cost figure behind a3d_query_order_work.  Results equal
 * a3d_bvh_udf's.  No counterpart in the reference. */
int a3d_bvh_udf_work(const void* blob, const float* P, int64_t K, const uint32_t* perm, float* dist,
                     int32_t* face, float* closest, float* uvw, int flags, uint32_t* work, a3d_stream_t stream);
/* a3d_bvh_udf_seeded: a3d_bvh_udf (wave walk) where seed[q] (float, indexed
 * like the outputs) is an upper bound on query q's distance: the walk starts
 * with it as its best distance and never enters a box farther away.  Results
 * equal a3d_bvh_udf's wherever the true distance is <= seed[q]; elsewhere
 * face[q] = -1 and dist[q] = seed[q] (the caller re-queries those without a
 * seed, or passes A3D_UDF_REDO to have them walked again in the same call,
 * which makes every result equal a3d_bvh_udf's).  flags: 0 or A3D_UDF_LANE,
 * plus A3D_UDF_REDO.  No counterpart in the reference: the surrogate
 * router's exact mode and the grid-seeded walk below.
 * a3d_udf_grid_seed: seed[K] for a3d_bvh_udf_seeded from a grid of exact
 * distances grid[(i (R+1) + j) (R+1) + k] at the points lo3 + (i, j, k) h
 * (float: lo + (float)i * h; i, j, k in [0, R], R <= 2048): for each query
 * the minimum over the 8 corners c of its cell (clamped into the grid) of
 * grid(c) + |q - c|, widened by 2^-12 relative and margin absolute (the
 * distance is 1-Lipschitz, so this bounds it up to FP32 rounding), and, when
 * grid_face (int32, each grid point's closest face; V float32[.,3] and F
 * int32[.,3] the mesh) is given, the distance from q to the nearest
 * corner's face by the walk's own formula.  perm (nullable): process the
 * queries in the walk's order (e.g. a3d_query_order's). */
int a3d_bvh_udf_seeded(const void* blob, const float* P, int64_t K, const uint32_t* perm, const float* seed,
                       float* dist, int32_t* face, float* closest, float* uvw, int flags, a3d_stream_t stream);
/* a3d_grid_points: the grid's points for a3d_udf_grid_seed / a3d_bvh_udf_grid
 * (R + 1 = side): P float32[side^3 * 3] = lo3 + (i, j, k) h (host lo3;
 * float: lo + (float)i * h) at index (i side + j) side + k, and perm
 * uint32[side^3], a walk order for a3d_bvh_udf in 4x4x4 bricks along a
 * Morton curve, so the grid's exact distances come from one plain walk with
 * no sort.  side a power of two in [4, 1024].  No counterpart in the
 * reference. */
int a3d_grid_points(const float* lo3, float h, int side, float* P, uint32_t* perm, a3d_stream_t stream);
int a3d_udf_grid_seed(const float* grid, const int32_t* grid_face, const float* V, const int32_t* F, int R,
                      const float* lo3, float h, const float* P, const uint32_t* perm, int64_t K, float margin,
                      float* seed, a3d_stream_t stream);
/* a3d_bvh_udf_grid: a3d_bvh_udf_seeded with A3D_UDF_REDO where each query's
 * bound is a3d_udf_grid_seed's, computed at the start of its walk (no seed
 * array): results equal a3d_bvh_udf's.  flags: as a3d_bvh_udf's except
 * A3D_UDF_SPLIT.  No counterpart in the reference. */
int a3d_bvh_udf_grid(const void* blob, const float* P, int64_t K, const uint32_t* perm, const float* grid,
                     const int32_t* grid_face, const float* V, const int32_t* F, int R, const float* lo3, float h,
                     float margin, float* dist, int32_t* face, float* closest, float* uvw, int flags,
                     a3d_stream_t stream);

/* ------------------------------------------------------------------ UDF adjoint
 * Replaces the autograd re-evaluation in MeshBVH._udf_with_grad
 * (core/mesh_bvh.py:685-700): dist = |P - C| per row with C constant;
 * backward grad_P = grad * (P - C) / dist, 0 where dist == 0.
 */
int a3d_point_dist(const float* P, const float* C, int64_t K, float* dist, a3d_stream_t stream);
int a3d_point_dist_backward(const float* P, const float* C, const float* dist, const float* grad,
                            int64_t K, float* grad_P, a3d_stream_t stream);

/* ------------------------------------------------------------------ SDF sign
 * Replaces MeshBVH._compute_sign (core/mesh_bvh.py:759-827): Baerentzen-
 * Aanaes pseudo-normal chosen by uvw (< 1e-4 per coordinate), sign of
 * dot(p - closest, n), 0 -> +1, face -1 -> +1.  Multiplies dist in place.
 * face_n float32[nF*3]; vertex_pn float32[nV*3], edge_pn float32[nE*3] and
 * he_edge int64[3*nF] are nullable (then the face normal is used).
 * F int32[nF*3] is needed only with vertex_pn.
 */
int a3d_sdf_sign(const float* P, int64_t K, const float* closest, const int32_t* face,
                 const float* uvw, const float* face_n, const float* vertex_pn,
                 const float* edge_pn, const int64_t* he_edge, const int32_t* F, float* dist,
                 a3d_stream_t stream);

/* ------------------------------------------------------------------ rays
 * Replace bvh_ray_intersect / bvh_ray_all_hit (bvh_kernels.cu:469-550,
 * 935-1022, wrappers :1068-1190, bound at :1247-1248) and the coalescing of
 * MeshBVH._signed_crossing_count (core/mesh_bvh.py:906-970).  O, D float32
 * [K*3] ray origins and directions (device), K < 2^31.  Per-face test: the
 * reference's Moller-Trumbore (bvh_kernels.cu:185-202, FP64 u/v range tests
 * as there).
 *
 * a3d_ray_nearest: hit uint8[K], t float32[K] (max_t when missed), face
 *   int32[K] (original index, -1), points float32[K*3] (o + d t; o when
 *   missed) and uvw float32[K*3] (Triangle::barycentric at the point; zeros
 *   when missed) -- points and uvw nullable.  Nearest = minimum t over all
 *   faces with t < max_t, ties to the lowest face index.
 * a3d_ray_all_count: counts int32[K] of the hits the reference's all-hit
 *   kernel keeps (0 < t < max_t, |dot(unit normal, d)| >= 1e-4).
 * a3d_ray_all_hits: with offsets int64[K] (exclusive prefix sum of the
 *   counts) writes each ray's hits from offsets[r] on, in walk order: ray
 *   index, face, and (nullable) t and sign (+1 / -1, sign of dot(n, d)).
 *   Sort the (ray, face) pairs with a3d_pairs_sort and call a3d_ray_eval for
 *   the canonical order.
 * a3d_ray_eval: t, sign, points, uvw (all nullable) of n (ray, face) pairs,
 *   from the mesh arrays V float32[nV*3], F int32[nF*3].
 * a3d_ray_crossings: signed crossing count out int32[K] from ray-major hit
 *   lists (offsets int64[K+1]); sorts each ray's (t, sign) in place by t.
 * a3d_ray_crossings_dirs: the same counts for N points P float32[N*3] along
 *   n_dirs directions dirs float32[n_dirs*3] (ray r = direction r / N from
 *   point r % N, N * n_dirs < 2^31) in one walk per ray, its hits kept
 *   sorted in registers: out int32[n_dirs*N]; a ray with more than 8 hits
 *   gets INT32_MIN and is counted by the two-pass path instead.  Replaces
 *   the per-direction loop of MeshBVH._sdf_parity (mesh_bvh.py:851-885).
 */
int a3d_ray_nearest(const void* blob, const float* O, const float* D, int64_t K, float max_t,
                    uint8_t* hit, float* t, int32_t* face, float* points, float* uvw, a3d_stream_t stream);
int a3d_ray_all_count(const void* blob, const float* O, const float* D, int64_t K, float max_t,
                      int32_t* counts, a3d_stream_t stream);
int a3d_ray_all_hits(const void* blob, const float* O, const float* D, int64_t K, float max_t,
                     const int64_t* offsets, int32_t* ray_idx, int32_t* face, float* t, int32_t* sign,
                     a3d_stream_t stream);
int a3d_ray_eval(const float* V, int64_t nV, const int32_t* F, int64_t nF, const float* O, const float* D,
                 int64_t n, const int32_t* ray_idx, const int32_t* face, float max_t, float* t,
                 int32_t* sign, float* points, float* uvw, a3d_stream_t stream);
int a3d_ray_crossings(float* t, int32_t* sign, const int64_t* offsets, int64_t K, float coalesce_t,
                      int32_t* out, a3d_stream_t stream);
int a3d_ray_crossings_dirs(const void* blob, const float* P, int64_t N, const float* dirs, int n_dirs, float max_t,
                           float coalesce_t, int32_t* out, a3d_stream_t stream);

/* ------------------------------------------------------------------ box-query tree
 * Replaces build_bvh (bvh_kernels.cu:860-925, bound at :1245) for the box
 * queries below: the reference's own binned-SAH tree (build_bvh_recursive
 * :667-854 — 16 bins, strict-minimum SAH split, std::partition, median
 * fallbacks by std::nth_element; leaf_max triangles per leaf, 4 in
 * kernels/bvh.py:69), built on the device with the same leaf sets, node
 * boxes and triangle order.  The reference's box query keeps a face when its
 * LEAF box overlaps the query box and the SAT passes, so its result depends
 * on this tree; the UDF and ray queries use the LBVH above instead (their
 * results do not).
 * a3d_sah_sizes: bytes of the blob and of the build workspace for nF faces.
 * a3d_sah_build: V float32[nV*3], F int32[nF*3] (device) -> blob (the BVH
 *   blob layout, without the UDF walk's extra boxes).  Synchronises the
 *   stream once per tree level above 64-triangle subtrees and at the end;
 *   *depth_out (host, nullable) = internal nodes on the longest path.  Fails
 *   with A3D_ERR_CAPACITY when the depth exceeds the box walks' stack (112;
 *   the reference's walk drops subtrees past 127).
 */
int a3d_sah_sizes(int64_t nF, int leaf_max, int64_t* blob_bytes, int64_t* ws_bytes);
int a3d_sah_build(const float* V, int64_t nV, const int32_t* F, int64_t nF, int leaf_max, void* blob,
                  int64_t blob_bytes, void* ws, int64_t ws_bytes, int32_t* depth_out, a3d_stream_t stream);

/* ------------------------------------------------------------------ AABB / SAT
 * Replaces bvh_aabb_intersect (bvh_kernels.cu:1192-1242 /
 * bvh_aabb_intersect_kernel :555-610, bound at :1249).  blob: from
 * a3d_sah_build.  A (box, face) pair hits when the box overlaps the face's
 * leaf box (inclusive) and the reference SAT (eps 1e-6) passes — the
 * reference's rule (:583-591).
 * a3d_bvh_aabb_hit:   hit uint8[K] only, first-hit early exit (mode 0).
 * a3d_bvh_aabb_count: counts int32[K] = pairs per box.
 * a3d_bvh_aabb_pairs: writes the pairs of box q at offsets[q] (int64[K],
 *                     exclusive prefix sum of counts), in traversal order
 *                     within each box; pair_a int32, pair_f int32.
 * a3d_pairs_sort:     sorts n pairs by (box, face) in place (the reference's
 *                     atomic append order is nondeterministic; this gives the
 *                     canonical order); workspace from a3d_pairs_sort_ws_bytes.
 */
int a3d_bvh_aabb_hit(const void* blob, const float* qmin, const float* qmax, int64_t K,
                     uint8_t* hit, a3d_stream_t stream);
int a3d_bvh_aabb_count(const void* blob, const float* qmin, const float* qmax, int64_t K,
                       int32_t* counts, a3d_stream_t stream);
int a3d_bvh_aabb_pairs(const void* blob, const float* qmin, const float* qmax, int64_t K,
                       const int64_t* offsets, int32_t* pair_a, int32_t* pair_f,
                       a3d_stream_t stream);
int a3d_pairs_sort_ws_bytes(int64_t n, int64_t* bytes);
int a3d_pairs_sort(int32_t* pair_a, int32_t* pair_f, int64_t n, void* ws, int64_t ws_bytes,
                   a3d_stream_t stream);

/* ------------------------------------------------------------------ polygon clip
 * Replaces sat_clip_polygon (cumtv_kernels.cu:1035-1100 /
 * sat_clip_polygon_kernel :193-460, bound at :1111), used by
 * MeshBVH.intersect_aabb modes 2 and 3 (core/mesh_bvh.py:381-479).
 * For each candidate k (box cand_a[k] of bmin/bmax float32[*,3], triangle
 * cand_t[k] of tris9 float32[M,9]): hit uint8[N], counts int32[N] (clipped
 * polygon vertices, <= 8), centroids float32[N,3], areas float32[N], and for
 * mode 2 poly float32[N,8,3] (zero padded).  mode: 0 hit only, 1 + centroid
 * and area, 2 + polygon.  Every output element is written.
 */
int a3d_sat_clip(const float* bmin, const float* bmax, const float* tris9, const int64_t* cand_a,
                 const int64_t* cand_t, int64_t N, int mode, float eps, uint8_t* hit, int32_t* counts,
                 float* poly, float* centroids, float* areas, a3d_stream_t stream);

/* ------------------------------------------------------------------ octree
 * Replaces the level loop of OctreeIndexer.octree_traverse
 * (grid/octree_indexer.py:470-509) and its per-level boxes
 * (cube_aabb_level :302-337, subdivide :191-216).  Cubes are packed uint64
 * keys i<<42 | j<<21 | k (level <= 21).
 * a3d_octree_ws_bytes: workspace for up to max_cubes candidate cubes.
 * a3d_octree_level (blob from a3d_sah_build; a cube is kept by the box
 *   query's rule above): candidates are all res^3 cubes of `level` in meshgrid('ij')
 *   order when parents == NULL, else the 8 children (CUBE_CORNERS order) of
 *   each parent key (level-1).  Writes the hit candidates, order preserved,
 *   to kept (capacity: the candidate count) and their number to n_kept
 *   (device int64).  bounds = host float[6] (b0 xyz, b1 xyz).
 * a3d_octree_unpack: keys -> int64[n*3] ijk.
 */
int a3d_octree_ws_bytes(int64_t max_cubes, int64_t* bytes);
int a3d_octree_level(const void* blob, const float* bounds_host, int level, const uint64_t* parents,
                     int64_t n_parents, uint64_t* kept, int64_t* n_kept, void* ws, int64_t ws_bytes,
                     a3d_stream_t stream);
int a3d_octree_unpack(const uint64_t* keys, int64_t n, int64_t* ijk, a3d_stream_t stream);

/* ------------------------------------------------------------------ flood fill
 * Replaces flood_fill (flood_fill_kernels.cu:117-175, kernel :21-87, bound
 * at :178) and the dense post-processing of flood_fill_3d_sparse
 * (kernels/flood_fill.py:118-258) and get_dam_neighborhood_signs
 * (apps/sparse_flood_fill.py:34-100).  The crop of D x H x W voxels (axis
 * order i, j, k) is held as A3D_FLOOD_NPLANES bit planes of plane_words
 * uint32 each (caller-owned `state`):
 *   OCC 0, REACHED 1, WATER 2 (mask 2), COLLISION 3 (mask 1), DAM 4 (mask -1
 *   incl. the extended dam), DRY 5 (mask -2), EXTENDED 6, NEIGHBORHOOD 7
 *   (k-ring 26-dilation of DAM), EXT_TIDE 8 (= NEIGHBORHOOD & (WATER |
 *   COLLISION)), DRY_ADJ 9 (= NEIGHBORHOOD & DRY), scratch 10.
 * a3d_flood_fill: occupancy from dam_ijk int64[N,3] (global coordinates,
 *   crop origin lo_host int64[3] on the host), or, with dam_ijk == NULL, from
 *   a prior a3d_flood_pack_dense; start (si, sj, sk) local; conn 6/18/26.
 *   Synchronises the stream.  totals_host[NPLANES] = set bits per plane;
 *   passes_host = global passes run, or -1 when the start voxel is occupied
 *   (every plane then empty, as the reference returns empty results).
 * a3d_flood_extract: coordinates (lo + local ijk) of one plane's set bits,
 *   lexicographic order (= torch.nonzero), int64[totals[plane], 3].
 * a3d_flood_dense_mask: the reference's int32 mask (-2/-1/1/2, dam without
 *   the extended dam) as a dense D x H x W array.
 * a3d_flood_dam_mask: uint8[N] = DAM bit at each dam_ijk row.
 */
#define A3D_FLOOD_NPLANES 11
/* a3d_ijk_bounds: out6 (device int64) = per-axis min (xyz) then max (xyz) of
 * ijk int64[N,3] (min = INT64_MAX, max = INT64_MIN when N = 0); the crop
 * bounds of flood_fill_3d_sparse (kernels/flood_fill.py:150-156). */
int a3d_ijk_bounds(const int64_t* ijk, int64_t N, int64_t* out6, a3d_stream_t stream);
int a3d_flood_sizes(int D, int H, int W, int64_t* plane_words, int64_t* ws_bytes);
int a3d_flood_pack_dense(const uint8_t* occ, int D, int H, int W, uint32_t* state, a3d_stream_t stream);
int a3d_flood_fill(const int64_t* dam_ijk, int64_t N, const int64_t* lo_host, int D, int H, int W,
                   int si, int sj, int sk, int conn, int k_ring, uint32_t* state, void* ws,
                   int64_t ws_bytes, int64_t* totals_host, int64_t* passes_host, a3d_stream_t stream);
int a3d_flood_extract(const uint32_t* state, int D, int H, int W, int plane, const int64_t* lo_host,
                      int64_t* out_ijk, void* ws, int64_t ws_bytes, a3d_stream_t stream);
int a3d_flood_dense_mask(const uint32_t* state, int D, int H, int W, int32_t* mask, a3d_stream_t stream);
int a3d_flood_dam_mask(const uint32_t* state, int D, int H, int W, const int64_t* dam_ijk, int64_t N,
                       const int64_t* lo_host, uint8_t* out, a3d_stream_t stream);

/* ------------------------------------------------------------------ DMC
 * Replaces SparseDiffDMC.forward in weighted_avg (training) mode and its
 * autograd backward (mesh_extractor/sparse_diffdmc.py:76-232 forward,
 * :234-261 weights and L_dev, :516-599 surface cubes / case ids / surface
 * edges, :601-780 dual vertices, :782-826 triangulation, :859-871
 * interpolation).  Inputs: vox int32[N,3] cube coordinates, ci int32[N,8]
 * corner ids (CUBE_CORNERS order) into sdf[M]; values float32 (dtype 0) or
 * float64 (dtype 1); optional deform [M,3], alpha [N,8], beta [N,12] (raw,
 * normalised inside as tanh(x) * weight_scale + 1), features [M,C].
 * a3d_dmc_ws_bytes: workspace for N cubes; holds the topology between calls.
 * a3d_dmc_topology: case ids (with the check_table flip), dual-vertex
 *   numbering, surface edges and quads, all integer and bit-exact with the
 *   reference.  tables_host = int8 dmc_table[256*4*7] | num_vd[256] |
 *   check_table[256*5].  Synchronises the stream; sizes_host[5] = surface
 *   cubes, dual vertices Vd, edge-group entries E, surface edges, quads Q.
 * a3d_dmc_faces: int64[4Q,3] training-mode faces (4 triangles per quad
 *   around a centre vertex Vd + q).
 * a3d_dmc_forward: verts [Vd+Q,3], L_dev [E], p_alpha [E,3], n_alpha [E,3],
 *   optional vfeat [Vd+Q,C].
 * a3d_dmc_backward: gradients of those outputs (each may be NULL = zero) to
 *   g_sdf [M] and, where non-NULL, g_deform [M,3], g_alpha [N,8], g_beta
 *   [N,12], g_feat [M,C] (all overwritten).  Follows an a3d_dmc_forward on
 *   the same workspace with the same inputs (which leaves each dual vertex,
 *   its weight sum and mean crossing distance there); the interpolation
 *   weights are recomputed, not stored.  gvf_scratch: [Vd,C] values, needed
 *   with g_vfeat.  scratch: scratch_bytes >= a3d_dmc_bwd_scratch_bytes(sizes,
 *   dtype) of caller memory (per-entry corner contributions and the sort
 *   buffers that replace atomics: every gradient except g_feat is summed in
 *   a fixed order, so repeated calls give identical bits).
 * a3d_dmc_bwd_scratch_bytes: the backward's scratch for a topology's sizes.
 * a3d_dmc_topology_get: int32 copies of the topology (kind 0 surface cube
 *   ids, 1 case ids, 2 entry -> dual vertex, 3 surface edges [.,2], 4 quads
 *   [.,4]).
 * QEF mode (the reference's eval-mode default; forward only):
 * a3d_dmc_corner_grads: replaces SparseDiffDMC._estimate_corner_gradients
 *   (sparse_diffdmc.py:450-513): per-corner least-squares SDF gradient over
 *   the unique edges of all N cubes, cgrad [M,3], each corner's sums taken
 *   in the reference's CPU order (no atomics).  scratch: caller memory of
 *   4*M int32.  Uses the workspace's sort buffers but none of the topology
 *   that a3d_dmc_topology leaves there; synchronises the stream once (twice
 *   for corner tables that are not a grid).
 * a3d_dmc_forward_qef: replaces the qef branch of _compute_vd and
 *   _solve_qef_scatter (sparse_diffdmc.py:617-655, 724-747, 289-389): the
 *   outputs of a3d_dmc_forward with dual vertices placed by the reference's
 *   regularised QEF (qef_reg_scale) and n_alpha from cgrad [M,3] (estimated
 *   or the caller's corner_gradients).  Quad centres are appended as in
 *   a3d_dmc_forward.  No backward.
 */
int a3d_dmc_ws_bytes(int64_t N, int64_t* bytes);
int a3d_dmc_topology(const int32_t* vox, const void* sdf, int dtype, double isovalue, const int32_t* ci, int64_t N,
                     int64_t M, int res, const int8_t* tables_host, void* ws, int64_t ws_bytes, int64_t* sizes_host,
                     a3d_stream_t stream);
int a3d_dmc_faces(const void* ws, int64_t N, const int64_t* sizes, int64_t* faces, a3d_stream_t stream);
int a3d_dmc_forward(void* ws, int64_t N, const int64_t* sizes, const int32_t* vox, const int32_t* ci,
                    const void* sdf, const void* deform, const void* alpha, const void* beta, const void* feat, int C,
                    int dtype, double isovalue, double weight_scale, int res, void* verts, void* L_dev,
                    void* p_alpha, void* n_alpha, void* vfeat, a3d_stream_t stream);
int a3d_dmc_backward(const void* ws, int64_t N, int64_t M, const int64_t* sizes, const int32_t* vox,
                     const int32_t* ci, const void* sdf, const void* deform, const void* alpha, const void* beta,
                     const void* feat, int C, const void* vfeat, int dtype, double isovalue, double weight_scale,
                     int res, const void* g_verts, const void* g_L_dev, const void* g_p_alpha, const void* g_n_alpha,
                     const void* g_vfeat, void* gvf_scratch, void* scratch, int64_t scratch_bytes, void* g_sdf,
                     void* g_deform, void* g_alpha, void* g_beta, void* g_feat, a3d_stream_t stream);
int a3d_dmc_bwd_scratch_bytes(const int64_t* sizes, int dtype, int64_t* bytes);
int a3d_dmc_topology_get(const void* ws, int64_t N, const int64_t* sizes, int kind, int32_t* out,
                         a3d_stream_t stream);
int a3d_dmc_corner_grads(void* ws, int64_t N, int64_t M, const int32_t* vox, const int32_t* ci, const void* sdf,
                         const void* deform, int dtype, double isovalue, int res, void* scratch, void* cgrad,
                         a3d_stream_t stream);
int a3d_dmc_forward_qef(void* ws, int64_t N, const int64_t* sizes, const int32_t* vox, const int32_t* ci,
                        const void* sdf, const void* deform, const void* alpha, const void* beta, const void* feat,
                        int C, int dtype, double isovalue, double weight_scale, int res, const void* cgrad,
                        double qef_reg_scale, void* verts, void* L_dev, void* p_alpha, void* n_alpha, void* vfeat,
                        a3d_stream_t stream);

/* ------------------------------------------------------------------ voxelizer
 * Replaces the candidate-cell SAT of Voxelizer.voxelize_surface /
 * voxelize_with_faces (apps/voxelizer.py:36-131), which materialise every
 * cell of the faces' bounding block and run MeshBVH.intersect_aabb on all of
 * them.  A cell (i,j,k) has CubeGrid.cube_aabb's box: min = float(ijk) *
 * cell + b0, max = min + cell (grid/cube_grid.py:164-203); b0, cell are
 * host float[3]; blob from a3d_sah_build (the box query's rule above).
 * a3d_vox_ws_bytes: workspace for a block of n_cells (< 2^31) cells.
 * a3d_vox_cells: cells of the block origin[3] + [0, dims[3]) (host int64)
 *   that intersect some face (the SAT test of intersect_aabb mode 0), in
 *   meshgrid('ij') order; *n_hit = their count, the list stays in the
 *   workspace.  Synchronises the stream.
 * a3d_vox_cells_get: writes those cells as int64 ijk [n_hit,3].
 * a3d_cells_sat_count / a3d_cells_sat_pairs: for given cells (int64 [n,3]),
 *   the SAT pair count per cell, then the (cell row, face id) pairs at
 *   offsets[n] (exclusive sums of the counts); a3d_pairs_sort puts them in
 *   (cell, face) order.
 */
int a3d_vox_ws_bytes(int64_t n_cells, int64_t* bytes);
int a3d_vox_cells(const void* blob, const float* b0, const float* cell, const int64_t* origin, const int64_t* dims,
                  void* ws, int64_t ws_bytes, int64_t* n_hit, a3d_stream_t stream);
int a3d_vox_cells_get(const void* ws, const int64_t* origin, const int64_t* dims, int64_t n_hit, int64_t* ijk,
                      a3d_stream_t stream);
int a3d_cells_sat_count(const void* blob, const int64_t* ijk, int64_t n, const float* b0, const float* cell,
                        int32_t* counts, a3d_stream_t stream);
int a3d_cells_sat_pairs(const void* blob, const int64_t* ijk, int64_t n, const float* b0, const float* cell,
                        const int64_t* offsets, int32_t* pair_a, int32_t* pair_f, a3d_stream_t stream);

/* ------------------------------------------------------------------ cells
 * Voxel-to-dual-graph glue between the flood fill / octree and the SDF /
 * DMC steps.  Cells are int32 [n,3] (i,j,k) in [0, resolution)^3; an input
 * outside that range is an A3D_ERR_ARG (the calls that synchronise report it).
 * method: 1 = radix sort of flat keys, 2 = bitmaps of the cell and vertex
 * sets ranked by popcount scans (resolution <= 2048), 0 = choose (bitmaps
 * when resolution <= 2048 and resolution^3 <= 4096 n).  Same results.
 * a3d_cells_ws_bytes: workspace for up to n cells (n <= 2^27) on that grid.
 * a3d_unique_cells: replaces the torch.unique(cat(dam, collision,
 *   dry_adjacent), dim=0) of construct_surface_dual_graph
 *   (apps/sparse_flood_fill.py:254-257): sorted (lexicographic) unique cells
 *   into out_ijk [n,3]; *n_out = their count.  Synchronises the stream.
 * a3d_cell_vertices: the unique corner vertices of n cells (duplicates
 *   allowed) in lexicographic order and cube_vert int64 [n,8], each cell
 *   corner's rank among them (CUBE_CORNERS order): replaces the vertex
 *   unique + searchsorted of construct_surface_dual_graph (:262-271) and the
 *   unique/inverse of CubeGrid.voxel_unique_vertices (grid/cube_grid.py:
 *   237-260).  *n_vert = vertex count; the vertices stay in the workspace.
 *   Synchronises the stream.
 * a3d_cell_vertices_get: writes those n_vert vertices (any output may be
 *   NULL): vkey int64 flat index in the (resolution+1)^3 vertex grid, vijk
 *   int64 [V,3], world float32 [V,3] = ijk * cell + lo (host float[3] each,
 *   two float32 roundings, as sparse_flood_fill.py:287-289).  Same n and
 *   method as the a3d_cell_vertices call.
 * a3d_cell_neighbors: adj int64 [n,6] for sorted unique cells: index of the
 *   cell at offset +x,-x,+y,-y,+z,-z, else -1, looked up by the unchecked
 *   flat key as the reference does (:273-283).  The bitmap method needs ws.
 * a3d_wet_ws_bytes / a3d_wet_mask: replaces get_dam_wet_faces / _edges /
 *   _corners (:169-227).  kind 0 faces -> out uint8 [n,6], 1 edges -> [n,12]
 *   (the reference's EDGE_NEIGHBOR_OFFSETS), 2 corners -> [n,8]: whether a
 *   neighbour (in range) is a collision voxel.  bitmap: caller memory of
 *   a3d_wet_ws_bytes bytes (resolution^3 bits), overwritten.  Collision
 *   voxels outside the grid are ignored.
 */
int a3d_cells_ws_bytes(int64_t n, int resolution, int method, int64_t* bytes);
int a3d_unique_cells(const int32_t* ijk, int64_t n, int resolution, int method, void* ws, int64_t ws_bytes,
                     int32_t* out_ijk, int64_t* n_out, a3d_stream_t stream);
int a3d_cell_vertices(const int32_t* ijk, int64_t n, int resolution, int method, void* ws, int64_t ws_bytes,
                      int64_t* cube_vert, int64_t* n_vert, a3d_stream_t stream);
int a3d_cell_vertices_get(const void* ws, int64_t n, int resolution, int method, int64_t n_vert, int64_t* vkey,
                          int64_t* vijk, float* world, const float* lo, const float* cell, a3d_stream_t stream);
int a3d_cell_neighbors(const int32_t* cells, int64_t n, int resolution, int method, void* ws, int64_t ws_bytes,
                       int64_t* adj, a3d_stream_t stream);
int a3d_wet_ws_bytes(int resolution, int64_t* bytes);
int a3d_wet_mask(int kind, const int32_t* dam, int64_t n, const int32_t* coll, int64_t c, int resolution,
                 void* bitmap, int64_t bitmap_bytes, uint8_t* out, a3d_stream_t stream);

/* ------------------------------------------------------------------ surrogate
 * Learned UDF surrogate: no counterpart in the reference.  A multiresolution
 * hash grid (16 levels, 2^19 entries of 2 features per level, resolutions
 * 16..2048 over the box [lo, hi], levels with (N+1)^3 <= 2^19 indexed
 * densely) feeding an MLP 32 -> 64 -> 64 -> 1 (ReLU; bf16 MFMA operands, fp32
 * accumulation).  `params` (a3d_sg_sizes) holds what inference reads: meta,
 * the fp16 table, bf16 W1 [64][32] and W2 [64][64], fp32 w3 [64] and b3.
 * `state` holds the fp32 master copies, Adam moments, gradients and
 * per-wave partial gradients used by training.
 * a3d_sg_init: domain and level layout, table ~ U(-1e-4, 1e-4), weights ~
 *   U(+-sqrt(6 / (fan_in + fan_out))), b3 = 0; clears the state.
 * a3d_sg_sample: n training points, a fraction near_frac within `radius` of
 *   the mesh surface, the rest uniform in [lo, hi].
 * a3d_sg_grad: gradients of mean((net(P) - target)^2) over a batch of B:
 *   the table gradient accumulates in the state; the MLP gradient vector
 *   (W1, W2, w3, b3 flattened, then the batch loss; 6272 floats) is written
 *   to the state and, if grad_out != NULL, copied there.
 * a3d_sg_table_grad: copy of the accumulated table gradient (16*2^19*2 floats).
 * a3d_sg_apply: one Adam step (bias-corrected for step t >= 1) on both, sparse
 *   on the table (entries with zero gradient untouched); clears the table
 *   gradient and refreshes the fp16 / bf16 copies in params.
 * a3d_sg_infer: out[q] = network(P[q]) (perm: optional query order).
 * a3d_sg_udf: the router.  Every query is answered by the network; those
 *   predicted below `band` are re-answered exactly by the BVH wave walk
 *   (a3d_bvh_udf results); the others get face -1, closest = the query
 *   point, uvw 0.  Writes the exact-query count to *n_near_host
 *   (synchronises the stream).  Workspace: a3d_sg_route_ws_bytes.
 */
int a3d_sg_sizes(int64_t* params_bytes, int64_t* state_bytes);
int a3d_sg_init(void* params, void* state, const float* lo_host, const float* hi_host, uint64_t seed,
                a3d_stream_t stream);
int a3d_sg_sample(const float* V, const int32_t* F, int64_t nF, int64_t n, float near_frac, float radius,
                  const float* lo_host, const float* hi_host, uint64_t seed, float* P, a3d_stream_t stream);
int a3d_sg_grad(const void* params, void* state, const float* P, const float* target, int64_t B, float* grad_out,
                a3d_stream_t stream);
int a3d_sg_table_grad(const void* state, float* out, a3d_stream_t stream);
int a3d_sg_apply(void* params, void* state, float lr, float beta1, float beta2, float eps, int64_t t,
                 a3d_stream_t stream);
int a3d_sg_infer(const void* params, const float* P, int64_t K, const uint32_t* perm, float* out,
                 a3d_stream_t stream);
int a3d_sg_route_ws_bytes(int64_t K, int64_t* bytes);
int a3d_sg_udf(const void* params, const void* blob, const float* P, int64_t K, const uint32_t* perm, float band,
               float* dist, int32_t* face, float* closest, float* uvw, void* ws, int64_t ws_bytes,
               int64_t* n_near_host, a3d_stream_t stream);

/* ------------------------------------------------------------------ batches
 * Closest-point queries over many meshes of uneven size in one persistent
 * launch.  No counterpart in the reference (one mesh per MeshBVH call).
 * blobs: device array of n_meshes BVH blob pointers; P: all queries
 * concatenated, mesh m owning [q_off[m], q_off[m+1]) (q_off device int64).
 * a3d_batch_order: perm (uint32 [K]) = queries sorted by mesh, then by Morton
 *   code in the mesh's box (8 bits per axis up to 256 meshes, 10 beyond;
 *   ties keep their input order).  Workspace: a3d_batch_order_ws_bytes.
 * a3d_batch_udf: tiles = int64 pairs (start << 8 | count, mesh), count <= 64,
 *   covering perm[start, start + count); a persistent grid of `blocks` blocks
 *   of `waves_per_block` wavefronts takes tiles from *counter (device uint32,
 *   reset by the call) in table order.  Results equal a3d_bvh_udf's on each
 *   mesh, written at the queries' positions.  Sync-free: capturable.
 *   bounds (nullable): device array of one a3d_grid_bound per mesh; each
 *   walk then starts bounded by its mesh's grid (as a3d_bvh_udf_grid) and a
 *   second pass walks the queries whose bound fell short (same results).
 * a3d_graph_begin / end / launch / destroy: capture the work queued on a
 *   stream into an instantiated HIP graph (opaque handle) and replay it.
 */
int a3d_batch_order_ws_bytes(int64_t K, int64_t* bytes);
int a3d_batch_order(const void* const* blobs, int n_meshes, const float* P, int64_t K, const int64_t* q_off,
                    uint32_t* perm, void* ws, int64_t ws_bytes, a3d_stream_t stream);
/* One mesh's distance-bound grid for a3d_batch_udf (64 bytes): the grid
 * of exact distances and closest faces of a3d_udf_grid_seed, the mesh's
 * vertices and faces, R, lo, h and the absolute margin. */
typedef struct a3d_grid_bound {
  const float* grid;
  const int32_t* grid_face;
  const float* V;
  const int32_t* F;
  int32_t R;
  float lo[3];
  float h;
  float margin;
  int32_t pad[2];
} a3d_grid_bound;
int a3d_batch_udf(const void* const* blobs, const float* P, const uint32_t* perm, const int64_t* tiles,
                  int64_t n_tiles, unsigned int* counter, int waves_per_block, int blocks, float* dist, int32_t* face,
                  float* closest, float* uvw, const void* bounds, a3d_stream_t stream);
int a3d_graph_begin(a3d_stream_t stream);
int a3d_graph_end(a3d_stream_t stream, int64_t* handle);
int a3d_graph_launch(int64_t handle, a3d_stream_t stream);
int a3d_graph_destroy(int64_t handle);

/* a3d_calib_valu: counter calibration, no counterpart in the reference.
 * Launches n threads (n a positive multiple of 256; out: device float [n])
 * that each run `iters` iterations of 8 independent FP32 fused
 * multiply-adds and write one sum; with the VALU and cycle counters this
 * gives the chip's issue rate of plain VALU instructions, the peak the
 * walks' VALU rates are compared with. */
int a3d_calib_valu(float* out, int64_t n, int iters, a3d_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* ATOM3D_AMD_H */
