// BVH queries for gfx950: closest point (UDF), SDF sign, AABB/SAT overlap
// walks and the octree level step.
//
// Traversal design (replaces the per-thread 128-entry private stacks of the
// reference, Atom3D atom3d/kernels/bvh_kernels.cu:363-463 and :555-610):
//   - one 64-byte node fetch yields both children's boxes, so each visited
//     node costs 4 x 16-byte loads and no dependent fetch;
//   - the traversal stack lives in LDS (kLds entries x 256 lanes, laid out
//     lane-minor so a wave's pushes hit 64 distinct banks), spilling to a
//     private array only past kLds entries (the tree depth is bounded by 97:
//     every level adds at least one bit to the Morton common prefix);
//   - blocks are remapped so each XCD walks a contiguous run of queries.
//
// Closest-face definition (tree independent): the face minimising the
// reference's FP32 squared distance, ties to the lowest original face id.
// Nodes are pruned only when their box distance exceeds the best distance by
// more than an error slack that covers (a) FP32 rounding of the box and
// triangle distances relative to the coordinates' magnitude and (b) the
// plane-branch error of thin triangles (folded into the leaf boxes at build
// time), so a face the reference formula would rank first is never skipped.
#include "a3d_prim.hpp"

#include "a3d_common.hpp"

namespace a3d {
namespace {

constexpr int kBlk = 256;
constexpr int kLds = 16;  // LDS stack entries per lane
constexpr int kOvf = kStackTotal - kLds;

// ---------------------------------------------------------------- AABB walks
// Box queries run on the reference-layout SAH tree (a3d_sah_build): a leaf
// is entered only when its box overlaps the query box (tested with the
// parent's record), then each of its triangles is kept when it passes the
// SAT test — the reference's rule (bvh_kernels.cu:583-591), so the kept set
// is the reference's, leaf grouping included.
// MODE 0: hit byte, stop at the first hit.  MODE 1: pair count.  MODE 2: write
// pairs at offsets[q].
template <int MODE>
__device__ __forceinline__ int32_t box_walk(const Node64* __restrict__ nodes, const Tri48* __restrict__ tris,
                                            v3 lo, v3 hi, int32_t* s_ref, int tid, int32_t qid,
                                            int32_t* __restrict__ pa, int32_t* __restrict__ pf, int64_t off) {
  int32_t ovf[kOvf];
  int sp = 0;
  int32_t cur = 0;
  int32_t cnt = 0;
  while (true) {
    if (cur >= 0) {
      const Node64* nd = nodes + cur;
      float4 a = nd->a, b = nd->b, c = nd->c;
      int4 d = nd->d;
      bool tl = box_overlap(a.x, a.y, a.z, a.w, b.x, b.y, lo, hi);
      bool tr = box_overlap(b.z, b.w, c.x, c.y, c.z, c.w, lo, hi);
      if (tl && tr) {
        if (sp < kLds) s_ref[sp * kBlk + tid] = d.y;
        else if (sp < kStackTotal) ovf[sp - kLds] = d.y;
        sp = sp < kStackTotal ? sp + 1 : sp;
        cur = d.x;
        continue;
      }
      if (tl) { cur = d.x; continue; }
      if (tr) { cur = d.y; continue; }
    } else {
      const int32_t s0 = leaf_start(cur), n = leaf_count(cur);
      for (int32_t i = s0; i < s0 + n; ++i) {
        v3 ta, tb, tc;
        int32_t id;
        load_tri(tris, i, ta, tb, tc, id);
        if (tri_aabb_sat(ta, tb, tc, lo, hi)) {
          if (MODE == 0) return 1;
          if (MODE == 2) {
            pa[off + cnt] = qid;
            pf[off + cnt] = id;
          }
          ++cnt;
        }
      }
    }
    if (sp == 0) break;
    --sp;
    cur = sp < kLds ? s_ref[sp * kBlk + tid] : ovf[sp - kLds];
  }
  return cnt;
}

template <int MODE>
__global__ __launch_bounds__(kBlk) void k_aabb(const void* __restrict__ blob, const float* __restrict__ qmin,
                                                const float* __restrict__ qmax, int64_t K, uint8_t* __restrict__ hit,
                                                int32_t* __restrict__ counts, const int64_t* __restrict__ offsets,
                                                int32_t* __restrict__ pa, int32_t* __restrict__ pf) {
  __shared__ int32_t s_ref[kLds * kBlk];
  const int tid = threadIdx.x;
  const int64_t q = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlk + tid;
  if (q >= K) return;
  v3 lo = mk(qmin[q * 3], qmin[q * 3 + 1], qmin[q * 3 + 2]);
  v3 hi = mk(qmax[q * 3], qmax[q * 3 + 1], qmax[q * 3 + 2]);
  int64_t off = MODE == 2 ? offsets[q] : 0;
  int32_t c = box_walk<MODE>(nodes_of(blob), tris_of(blob), lo, hi, s_ref, tid, (int32_t)q, pa, pf, off);
  if (MODE == 0) hit[q] = c ? 1 : 0;
  if (MODE == 1) counts[q] = c;
}

// ---------------------------------------------------------------- voxelizer
// Cells of an axis-aligned block of a regular grid, tested against the mesh
// with the SAT walk (Voxelizer.voxelize_surface / voxelize_with_faces,
// Atom3D atom3d/apps/voxelizer.py:36-131).  A cell's box is
// CubeGrid.cube_aabb's: min = float(ijk) * cell + b0 (two roundings), max =
// min + cell (grid/cube_grid.py:164-203).
struct VoxGeom {
  float b0[3];
  float cell[3];
  int64_t org[3];
  int64_t dims[3];
};

__device__ __forceinline__ void vox_box(const VoxGeom& g, int64_t i, int64_t j, int64_t k, v3& lo, v3& hi) {
  const float ijk[3] = {(float)i, (float)j, (float)k};
  float l[3], h[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float t = ijk[a] * g.cell[a];
    l[a] = t + g.b0[a];
    h[a] = l[a] + g.cell[a];
  }
  lo = mk(l[0], l[1], l[2]);
  hi = mk(h[0], h[1], h[2]);
}

// Candidate c of the block (meshgrid 'ij' order: k fastest) -> hit flag by
// the first-hit walk.
__global__ __launch_bounds__(kBlk) void k_vox_test(const void* __restrict__ blob, VoxGeom g, int64_t n,
                                                    uint8_t* __restrict__ flags) {
  __shared__ int32_t s_ref[kLds * kBlk];
  const int tid = threadIdx.x;
  const int64_t c = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlk + tid;
  if (c >= n) return;
  const int64_t i = g.org[0] + c / (g.dims[1] * g.dims[2]);
  const int64_t j = g.org[1] + (c / g.dims[2]) % g.dims[1];
  const int64_t k = g.org[2] + c % g.dims[2];
  v3 lo, hi;
  vox_box(g, i, j, k, lo, hi);
  flags[c] = box_walk<0>(nodes_of(blob), tris_of(blob), lo, hi, s_ref, tid, 0, nullptr, nullptr, 0) ? 1 : 0;
}

__global__ void k_vox_unravel(const int32_t* __restrict__ idx, int64_t n, VoxGeom g, int64_t* __restrict__ ijk) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= n) return;
  const int64_t c = idx[t];
  ijk[t * 3 + 0] = g.org[0] + c / (g.dims[1] * g.dims[2]);
  ijk[t * 3 + 1] = g.org[1] + (c / g.dims[2]) % g.dims[1];
  ijk[t * 3 + 2] = g.org[2] + c % g.dims[2];
}

// Given cells (int64 ijk): MODE 1 SAT pair count, MODE 2 pairs (cell row,
// face id) written at offsets[q].
template <int MODE>
__global__ __launch_bounds__(kBlk) void k_cells_sat(const void* __restrict__ blob, VoxGeom g,
                                                     const int64_t* __restrict__ ijk, int64_t n,
                                                     int32_t* __restrict__ counts, const int64_t* __restrict__ offsets,
                                                     int32_t* __restrict__ pa, int32_t* __restrict__ pf) {
  __shared__ int32_t s_ref[kLds * kBlk];
  const int tid = threadIdx.x;
  const int64_t q = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlk + tid;
  if (q >= n) return;
  v3 lo, hi;
  vox_box(g, ijk[q * 3], ijk[q * 3 + 1], ijk[q * 3 + 2], lo, hi);
  const int64_t off = MODE == 2 ? offsets[q] : 0;
  const int32_t c = box_walk<MODE>(nodes_of(blob), tris_of(blob), lo, hi, s_ref, tid, (int32_t)q, pa, pf, off);
  if (MODE == 1) counts[q] = c;
}

// ---------------------------------------------------------------- octree
struct LevelGeom {
  float b0[3];
  float cell[3];
  int64_t res;
};

__device__ __forceinline__ uint64_t pack_ijk(uint64_t i, uint64_t j, uint64_t k) { return (i << 42) | (j << 21) | k; }

// Candidate c -> cube key; box exactly as cube_aabb_level computes it:
// min = float(ijk) * cell + b0 (two roundings), max = min + cell.
__global__ __launch_bounds__(kBlk) void k_octree_test(const void* __restrict__ blob, LevelGeom g,
                                                       const uint64_t* __restrict__ parents, int64_t n_cand,
                                                       uint64_t* __restrict__ keys, uint8_t* __restrict__ flags) {
  __shared__ int32_t s_ref[kLds * kBlk];
  const int tid = threadIdx.x;
  const int64_t c = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlk + tid;
  if (c >= n_cand) return;
  uint64_t i, j, k;
  if (parents == nullptr) {
    i = (uint64_t)(c / (g.res * g.res));
    j = (uint64_t)((c / g.res) % g.res);
    k = (uint64_t)(c % g.res);
  } else {
    uint64_t pk = parents[c >> 3];
    int o = (int)(c & 7);
    i = ((pk >> 42) & 0x1fffff) * 2 + (o & 1);
    j = ((pk >> 21) & 0x1fffff) * 2 + ((o >> 1) & 1);
    k = (pk & 0x1fffff) * 2 + ((o >> 2) & 1);
  }
  float ijk[3] = {(float)i, (float)j, (float)k};
  float lo[3], hi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float t = ijk[a] * g.cell[a];
    lo[a] = t + g.b0[a];
    hi[a] = lo[a] + g.cell[a];
  }
  int32_t h = box_walk<0>(nodes_of(blob), tris_of(blob), mk(lo[0], lo[1], lo[2]), mk(hi[0], hi[1], hi[2]), s_ref,
                          tid, 0, nullptr, nullptr, 0);
  keys[c] = pack_ijk(i, j, k);
  flags[c] = h ? 1 : 0;
}

__global__ __launch_bounds__(kBlk) void k_unpack(const uint64_t* __restrict__ keys, int64_t n, int64_t* __restrict__ ijk) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= n) return;
  uint64_t key = keys[t];
  ijk[t * 3] = (int64_t)((key >> 42) & 0x1fffff);
  ijk[t * 3 + 1] = (int64_t)((key >> 21) & 0x1fffff);
  ijk[t * 3 + 2] = (int64_t)(key & 0x1fffff);
}

// ---------------------------------------------------------------- pair sort
__global__ __launch_bounds__(kBlk) void k_pairs_pack(const int32_t* __restrict__ a, const int32_t* __restrict__ f,
                                                      int64_t n, uint64_t* __restrict__ key) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= n) return;
  key[t] = ((uint64_t)(uint32_t)a[t] << 32) | (uint64_t)(uint32_t)f[t];
}
__global__ __launch_bounds__(kBlk) void k_pairs_unpack(const uint64_t* __restrict__ key, int64_t n,
                                                        int32_t* __restrict__ a, int32_t* __restrict__ f) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= n) return;
  uint64_t k = key[t];
  a[t] = (int32_t)(k >> 32);
  f[t] = (int32_t)(k & 0xffffffffu);
}

// ---------------------------------------------------------------- SDF sign
// MeshBVH._compute_sign (core/mesh_bvh.py:759-827).
__global__ __launch_bounds__(kBlk) void k_sdf_sign(const float* __restrict__ P, int64_t K,
                                                    const float* __restrict__ closest, const int32_t* __restrict__ face,
                                                    const float* __restrict__ uvw, const float* __restrict__ face_n,
                                                    const float* __restrict__ vertex_pn,
                                                    const float* __restrict__ edge_pn,
                                                    const int64_t* __restrict__ he_edge,
                                                    const int32_t* __restrict__ F, float* __restrict__ dist) {
  int64_t q = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (q >= K) return;
  int32_t f = face[q];
  if (f < 0) return;  // sign +1
  const float* n = face_n + (int64_t)f * 3;
  if (uvw && edge_pn && he_edge) {
    const float eps = 1e-4f;
    float b0 = uvw[q * 3], b1 = uvw[q * 3 + 1], b2 = uvw[q * 3 + 2];
    bool z0 = b0 < eps, z1 = b1 < eps, z2 = b2 < eps;
    int nz = (int)z0 + (int)z1 + (int)z2;
    if (nz == 2) {
      int dom = !z0 ? 0 : (!z1 ? 1 : 2);  // first coordinate not near zero
      int32_t vid = F[(int64_t)f * 3 + dom];
      n = vertex_pn + (int64_t)vid * 3;
    } else if (nz == 1) {
      int zi = z0 ? 0 : (z1 ? 1 : 2);
      int he_local = zi == 0 ? 1 : (zi == 1 ? 2 : 0);
      int64_t e = he_edge[(int64_t)f * 3 + he_local];
      n = edge_pn + e * 3;
    }
  }
  float dx = P[q * 3] - closest[q * 3];
  float dy = P[q * 3 + 1] - closest[q * 3 + 1];
  float dz = P[q * 3 + 2] - closest[q * 3 + 2];
  float dt = (dx * n[0] + dy * n[1]) + dz * n[2];
  if (dt < 0.0f) dist[q] = -dist[q];
}

// ---------------------------------------------------------------- UDF adjoint
// Differentiable distance |p - c| with c held constant (the reference
// re-evaluates (points - closest.detach()).norm(dim=1) under autograd,
// core/mesh_bvh.py:685-700).  Forward sums squares in the same order as a
// 3-wide row reduction; backward is g * (p - c) / d, and 0 where d == 0.
__global__ __launch_bounds__(kBlk) void k_pdist_fwd(const float* __restrict__ P, const float* __restrict__ C,
                                                     int64_t K, float* __restrict__ dist) {
  int64_t q = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (q >= K) return;
  float dx = P[q * 3] - C[q * 3], dy = P[q * 3 + 1] - C[q * 3 + 1], dz = P[q * 3 + 2] - C[q * 3 + 2];
  dist[q] = sqrtf((dx * dx + dy * dy) + dz * dz);
}

__global__ __launch_bounds__(kBlk) void k_pdist_bwd(const float* __restrict__ P, const float* __restrict__ C,
                                                     const float* __restrict__ dist, const float* __restrict__ g,
                                                     int64_t K, float* __restrict__ gP) {
  int64_t q = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (q >= K) return;
  float d = dist[q];
  float s = d == 0.0f ? 0.0f : g[q] / d;
  gP[q * 3] = (P[q * 3] - C[q * 3]) * s;
  gP[q * 3 + 1] = (P[q * 3 + 1] - C[q * 3 + 1]) * s;
  gP[q * 3 + 2] = (P[q * 3 + 2] - C[q * 3 + 2]) * s;
}

__global__ void k_check_blob(const void* blob, int32_t* ok) {
  if (threadIdx.x == 0) *ok = header(blob)->magic == kBvhMagic;
}

inline unsigned grid_for(int64_t n) { return (unsigned)div_up(n > 0 ? n : 1, kBlk); }

int check_blob(const void* blob, const char* who) {
  if (!blob) return fail(A3D_ERR_ARG, "%s: null BVH blob", who);
  return A3D_OK;
}

size_t octree_cub_bytes(int64_t n) {
  size_t b = 0;
  (void)rocprim::select(nullptr, b, (uint64_t*)nullptr, (uint8_t*)nullptr, (uint64_t*)nullptr,
                                (int64_t*)nullptr, (int)(n > 1 ? n : 2));
  return b;
}

size_t pairs_cub_bytes(int64_t n) {
  size_t b = 0;
  (void)rocprim::radix_sort_keys(nullptr, b, (uint64_t*)nullptr, (uint64_t*)nullptr, (int)(n > 1 ? n : 2));
  return b;
}

inline int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

}  // namespace
}  // namespace a3d

using namespace a3d;

extern "C" int a3d_sdf_sign(const float* P, int64_t K, const float* closest, const int32_t* face, const float* uvw,
                            const float* face_n, const float* vertex_pn, const float* edge_pn, const int64_t* he_edge,
                            const int32_t* F, float* dist, a3d_stream_t stream) {
  if (K < 0) return fail(A3D_ERR_ARG, "a3d_sdf_sign: K < 0");
  if (K == 0) return A3D_OK;
  if (!P || !closest || !face || !face_n || !dist) return fail(A3D_ERR_ARG, "a3d_sdf_sign: null input/output");
  if (edge_pn && he_edge && uvw && (!vertex_pn || !F))
    return fail(A3D_ERR_ARG, "a3d_sdf_sign: vertex_pn and F are required with edge pseudo-normals");
  hipLaunchKernelGGL(k_sdf_sign, dim3(grid_for(K)), dim3(kBlk), 0, (hipStream_t)stream, P, K, closest, face, uvw,
                     face_n, vertex_pn, edge_pn, he_edge, F, dist);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_bvh_aabb_hit(const void* blob, const float* qmin, const float* qmax, int64_t K, uint8_t* hit,
                                a3d_stream_t stream) {
  if (int rc = check_blob(blob, "a3d_bvh_aabb_hit")) return rc;
  if (K < 0) return fail(A3D_ERR_ARG, "a3d_bvh_aabb_hit: K < 0");
  if (K == 0) return A3D_OK;
  if (!qmin || !qmax || !hit) return fail(A3D_ERR_ARG, "a3d_bvh_aabb_hit: null input/output");
  hipLaunchKernelGGL(k_aabb<0>, dim3(grid_for(K)), dim3(kBlk), 0, (hipStream_t)stream, blob, qmin, qmax, K, hit,
                     (int32_t*)nullptr, (const int64_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_bvh_aabb_count(const void* blob, const float* qmin, const float* qmax, int64_t K, int32_t* counts,
                                  a3d_stream_t stream) {
  if (int rc = check_blob(blob, "a3d_bvh_aabb_count")) return rc;
  if (K < 0) return fail(A3D_ERR_ARG, "a3d_bvh_aabb_count: K < 0");
  if (K == 0) return A3D_OK;
  if (!qmin || !qmax || !counts) return fail(A3D_ERR_ARG, "a3d_bvh_aabb_count: null input/output");
  hipLaunchKernelGGL(k_aabb<1>, dim3(grid_for(K)), dim3(kBlk), 0, (hipStream_t)stream, blob, qmin, qmax, K,
                     (uint8_t*)nullptr, counts, (const int64_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_bvh_aabb_pairs(const void* blob, const float* qmin, const float* qmax, int64_t K,
                                  const int64_t* offsets, int32_t* pair_a, int32_t* pair_f, a3d_stream_t stream) {
  if (int rc = check_blob(blob, "a3d_bvh_aabb_pairs")) return rc;
  if (K < 0) return fail(A3D_ERR_ARG, "a3d_bvh_aabb_pairs: K < 0");
  if (K == 0) return A3D_OK;
  if (!qmin || !qmax || !offsets || !pair_a || !pair_f) return fail(A3D_ERR_ARG, "a3d_bvh_aabb_pairs: null input/output");
  hipLaunchKernelGGL(k_aabb<2>, dim3(grid_for(K)), dim3(kBlk), 0, (hipStream_t)stream, blob, qmin, qmax, K,
                     (uint8_t*)nullptr, (int32_t*)nullptr, offsets, pair_a, pair_f);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_pairs_sort_ws_bytes(int64_t n, int64_t* bytes) {
  if (n < 0 || !bytes) return fail(A3D_ERR_ARG, "a3d_pairs_sort_ws_bytes: bad argument");
  *bytes = 2 * al256(8 * (n > 0 ? n : 1)) + al256((int64_t)pairs_cub_bytes(n));
  return A3D_OK;
}

extern "C" int a3d_pairs_sort(int32_t* pair_a, int32_t* pair_f, int64_t n, void* ws, int64_t ws_bytes,
                              a3d_stream_t stream) {
  int64_t need = 0;
  if (int rc = a3d_pairs_sort_ws_bytes(n, &need)) return rc;
  if (n == 0) return A3D_OK;
  if (n > 0x7fffffff) return fail(A3D_ERR_ARG, "a3d_pairs_sort: n=%lld exceeds 2^31-1", (long long)n);
  if (!pair_a || !pair_f || !ws) return fail(A3D_ERR_ARG, "a3d_pairs_sort: null argument");
  if (ws_bytes < need) return fail(A3D_ERR_CAPACITY, "a3d_pairs_sort: workspace %lld < %lld", (long long)ws_bytes, (long long)need);
  hipStream_t s = (hipStream_t)stream;
  char* w = (char*)ws;
  uint64_t* k0 = (uint64_t*)w;
  uint64_t* k1 = (uint64_t*)(w + al256(8 * n));
  void* tmp = w + 2 * al256(8 * n);
  size_t tb = pairs_cub_bytes(n);
  hipLaunchKernelGGL(k_pairs_pack, dim3(grid_for(n)), dim3(kBlk), 0, s, pair_a, pair_f, n, k0);
  A3D_LAUNCH_CHECK();
  A3D_HIP_CHECK(rocprim::radix_sort_keys(tmp, tb, k0, k1, (int)n, 0, 64, s));
  hipLaunchKernelGGL(k_pairs_unpack, dim3(grid_for(n)), dim3(kBlk), 0, s, k1, n, pair_a, pair_f);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_octree_ws_bytes(int64_t max_cubes, int64_t* bytes) {
  if (max_cubes < 0 || !bytes) return fail(A3D_ERR_ARG, "a3d_octree_ws_bytes: bad argument");
  int64_t n = max_cubes > 0 ? max_cubes : 1;
  *bytes = al256(8 * n) + al256(n) + al256((int64_t)octree_cub_bytes(n));
  return A3D_OK;
}

extern "C" int a3d_octree_level(const void* blob, const float* bounds_host, int level, const uint64_t* parents,
                                int64_t n_parents, uint64_t* kept, int64_t* n_kept, void* ws, int64_t ws_bytes,
                                a3d_stream_t stream) {
  if (int rc = check_blob(blob, "a3d_octree_level")) return rc;
  if (!bounds_host || !kept || !n_kept || !ws) return fail(A3D_ERR_ARG, "a3d_octree_level: null argument");
  if (level < 0 || level > 21) return fail(A3D_ERR_ARG, "a3d_octree_level: level=%d not in [0, 21]", level);
  if (parents && level == 0) return fail(A3D_ERR_ARG, "a3d_octree_level: level 0 has no parents");
  if (n_parents < 0) return fail(A3D_ERR_ARG, "a3d_octree_level: n_parents < 0");
  int64_t res = int64_t(1) << level;
  int64_t n_cand = parents ? n_parents * 8 : res * res * res;
  if (n_cand > 0x7fffffff) return fail(A3D_ERR_ARG, "a3d_octree_level: %lld candidates exceed 2^31-1", (long long)n_cand);
  int64_t need = 0;
  a3d_octree_ws_bytes(n_cand, &need);
  if (ws_bytes < need) return fail(A3D_ERR_CAPACITY, "a3d_octree_level: workspace %lld < %lld", (long long)ws_bytes, (long long)need);
  hipStream_t s = (hipStream_t)stream;
  LevelGeom g;
  for (int a = 0; a < 3; ++a) {
    g.b0[a] = bounds_host[a];
    g.cell[a] = (bounds_host[3 + a] - bounds_host[a]) / (float)res;
  }
  g.res = res;
  char* w = (char*)ws;
  int64_t nn = n_cand > 0 ? n_cand : 1;
  uint64_t* keys = (uint64_t*)w;
  uint8_t* flags = (uint8_t*)(w + al256(8 * nn));
  void* tmp = w + al256(8 * nn) + al256(nn);
  if (n_cand == 0) {
    A3D_HIP_CHECK(hipMemcpyAsync(n_kept, &n_cand, sizeof(int64_t), hipMemcpyHostToDevice, s));
    A3D_HIP_CHECK(hipStreamSynchronize(s));
    return A3D_OK;
  }
  hipLaunchKernelGGL(k_octree_test, dim3(grid_for(n_cand)), dim3(kBlk), 0, s, blob, g, parents, n_cand, keys, flags);
  A3D_LAUNCH_CHECK();
  size_t tb = octree_cub_bytes(n_cand);
  A3D_HIP_CHECK(rocprim::select(tmp, tb, keys, flags, kept, n_kept, (int)n_cand, s));
  return A3D_OK;
}

extern "C" int a3d_octree_unpack(const uint64_t* keys, int64_t n, int64_t* ijk, a3d_stream_t stream) {
  if (n < 0) return fail(A3D_ERR_ARG, "a3d_octree_unpack: n < 0");
  if (n == 0) return A3D_OK;
  if (!keys || !ijk) return fail(A3D_ERR_ARG, "a3d_octree_unpack: null argument");
  hipLaunchKernelGGL(k_unpack, dim3(grid_for(n)), dim3(kBlk), 0, (hipStream_t)stream, keys, n, ijk);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_point_dist(const float* P, const float* C, int64_t K, float* dist, a3d_stream_t stream) {
  if (K < 0) return fail(A3D_ERR_ARG, "a3d_point_dist: K < 0");
  if (K == 0) return A3D_OK;
  if (!P || !C || !dist) return fail(A3D_ERR_ARG, "a3d_point_dist: null argument");
  hipLaunchKernelGGL(k_pdist_fwd, dim3(grid_for(K)), dim3(kBlk), 0, (hipStream_t)stream, P, C, K, dist);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_point_dist_backward(const float* P, const float* C, const float* dist, const float* grad,
                                       int64_t K, float* grad_P, a3d_stream_t stream) {
  if (K < 0) return fail(A3D_ERR_ARG, "a3d_point_dist_backward: K < 0");
  if (K == 0) return A3D_OK;
  if (!P || !C || !dist || !grad || !grad_P) return fail(A3D_ERR_ARG, "a3d_point_dist_backward: null argument");
  hipLaunchKernelGGL(k_pdist_bwd, dim3(grid_for(K)), dim3(kBlk), 0, (hipStream_t)stream, P, C, dist, grad, K, grad_P);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

// ---------------------------------------------------------------- voxelizer host
namespace {
size_t vox_cub_bytes(int64_t n) {
  size_t t = 0;
  rocprim::counting_iterator<int32_t> it(0);
  (void)rocprim::select(nullptr, t, it, (uint8_t*)nullptr, (int32_t*)nullptr, (int64_t*)nullptr,
                                      (int)(n > 0 ? n : 1));
  return t;
}

int vox_geom(const float* b0, const float* cell, const int64_t* org, const int64_t* dims, VoxGeom& g,
             const char* who) {
  if (!b0 || !cell) return fail(A3D_ERR_ARG, "%s: null b0 or cell", who);
  for (int a = 0; a < 3; ++a) {
    g.b0[a] = b0[a];
    g.cell[a] = cell[a];
    g.org[a] = org ? org[a] : 0;
    g.dims[a] = dims ? dims[a] : 1;
    if (g.dims[a] < 0 || g.org[a] < 0) return fail(A3D_ERR_ARG, "%s: negative block origin or size", who);
  }
  return A3D_OK;
}
}  // namespace

extern "C" int a3d_vox_ws_bytes(int64_t n_cells, int64_t* bytes) {
  if (n_cells < 0 || n_cells > 0x7fffffff || !bytes)
    return fail(A3D_ERR_ARG, "a3d_vox_ws_bytes: n_cells=%lld not in [0, 2^31-1]", (long long)n_cells);
  const int64_t n = n_cells > 0 ? n_cells : 1;
  *bytes = al256(n) + al256(4 * n) + al256(8) + al256((int64_t)vox_cub_bytes(n));
  return A3D_OK;
}

extern "C" int a3d_vox_cells(const void* blob, const float* b0, const float* cell, const int64_t* origin,
                             const int64_t* dims, void* ws, int64_t ws_bytes, int64_t* n_hit, a3d_stream_t stream) {
  if (int rc = check_blob(blob, "a3d_vox_cells")) return rc;
  if (!origin || !dims || !n_hit || !ws) return fail(A3D_ERR_ARG, "a3d_vox_cells: null argument");
  VoxGeom g;
  if (int rc = vox_geom(b0, cell, origin, dims, g, "a3d_vox_cells")) return rc;
  const int64_t n = g.dims[0] * g.dims[1] * g.dims[2];
  int64_t need = 0;
  if (int rc = a3d_vox_ws_bytes(n, &need)) return rc;
  if (ws_bytes < need) return fail(A3D_ERR_CAPACITY, "a3d_vox_cells: workspace %lld < %lld", (long long)ws_bytes,
                                   (long long)need);
  *n_hit = 0;
  if (n == 0) return A3D_OK;
  hipStream_t s = (hipStream_t)stream;
  char* w = (char*)ws;
  uint8_t* flags = (uint8_t*)w;
  int32_t* idx = (int32_t*)(w + al256(n));
  int64_t* cnt = (int64_t*)(w + al256(n) + al256(4 * n));
  void* tmp = w + al256(n) + al256(4 * n) + al256(8);
  hipLaunchKernelGGL(k_vox_test, dim3(grid_for(n)), dim3(kBlk), 0, s, blob, g, n, flags);
  A3D_LAUNCH_CHECK();
  size_t tb = vox_cub_bytes(n);
  rocprim::counting_iterator<int32_t> it(0);
  A3D_HIP_CHECK(rocprim::select(tmp, tb, it, flags, idx, cnt, (int)n, s));
  A3D_HIP_CHECK(hipMemcpyAsync(n_hit, cnt, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  A3D_HIP_CHECK(hipStreamSynchronize(s));
  return A3D_OK;
}

extern "C" int a3d_vox_cells_get(const void* ws, const int64_t* origin, const int64_t* dims, int64_t n_hit,
                                 int64_t* ijk, a3d_stream_t stream) {
  if (!ws || !origin || !dims || n_hit < 0 || (n_hit > 0 && !ijk))
    return fail(A3D_ERR_ARG, "a3d_vox_cells_get: bad arguments");
  const float z[3] = {0.f, 0.f, 0.f};
  VoxGeom g;
  if (int rc = vox_geom(z, z, origin, dims, g, "a3d_vox_cells_get")) return rc;
  const int64_t n = g.dims[0] * g.dims[1] * g.dims[2];
  if (n_hit > n) return fail(A3D_ERR_ARG, "a3d_vox_cells_get: n_hit %lld > cells %lld", (long long)n_hit, (long long)n);
  if (n_hit == 0) return A3D_OK;
  const int32_t* idx = (const int32_t*)((const char*)ws + al256(n));
  hipLaunchKernelGGL(k_vox_unravel, dim3(grid_for(n_hit)), dim3(kBlk), 0, (hipStream_t)stream, idx, n_hit, g, ijk);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_cells_sat_count(const void* blob, const int64_t* ijk, int64_t n, const float* b0, const float* cell,
                                   int32_t* counts, a3d_stream_t stream) {
  if (int rc = check_blob(blob, "a3d_cells_sat_count")) return rc;
  if (n < 0 || n > 0x7fffffff || (n > 0 && (!ijk || !counts)))
    return fail(A3D_ERR_ARG, "a3d_cells_sat_count: bad arguments");
  VoxGeom g;
  if (int rc = vox_geom(b0, cell, nullptr, nullptr, g, "a3d_cells_sat_count")) return rc;
  if (n == 0) return A3D_OK;
  hipLaunchKernelGGL(k_cells_sat<1>, dim3(grid_for(n)), dim3(kBlk), 0, (hipStream_t)stream, blob, g, ijk, n, counts,
                     nullptr, nullptr, nullptr);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_cells_sat_pairs(const void* blob, const int64_t* ijk, int64_t n, const float* b0, const float* cell,
                                   const int64_t* offsets, int32_t* pair_a, int32_t* pair_f, a3d_stream_t stream) {
  if (int rc = check_blob(blob, "a3d_cells_sat_pairs")) return rc;
  if (n < 0 || n > 0x7fffffff || (n > 0 && (!ijk || !offsets || !pair_a || !pair_f)))
    return fail(A3D_ERR_ARG, "a3d_cells_sat_pairs: bad arguments");
  VoxGeom g;
  if (int rc = vox_geom(b0, cell, nullptr, nullptr, g, "a3d_cells_sat_pairs")) return rc;
  if (n == 0) return A3D_OK;
  hipLaunchKernelGGL(k_cells_sat<2>, dim3(grid_for(n)), dim3(kBlk), 0, (hipStream_t)stream, blob, g, ijk, n, nullptr,
                     offsets, pair_a, pair_f);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}
