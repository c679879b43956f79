// Device BVH build for gfx950: Morton-ordered linear BVH (Karras 2012,
// "Maximizing Parallelism in the Construction of BVHs, Octrees, and k-d
// Trees"), collapse of subtrees of <= leaf_max triangles into leaves, box
// fitting of the remaining internal nodes in frontier rounds, optional SAH
// treelet restructuring of those nodes (Karras & Aila 2013, "Fast Parallel
// Construction of High-Quality Bounding Volume Hierarchies"), and emission of
// 64-byte two-child nodes.
//
// Replaces the reference's single-threaded host SAH build
// (Atom3D atom3d/kernels/bvh_kernels.cu:616-925), which copies the mesh to
// the host and back.  Query results never depend on the tree shape: the
// traversals return the same faces for any valid tree.
//
// Bottom-up passes run as frontier rounds, one kernel launch per tree level:
// a node enters the frontier of round r+1 when the second of its children
// finishes in round r (relaxed device-scope arrival counter), and the node's
// children are read only in a later launch.  No box crosses workgroups
// inside one launch, so no device-scope fence (an L2 write-back on a
// multi-XCD part) is needed.
#include "a3d_prim.hpp"

#include <cstddef>
#include <stdarg.h>
#include <stdio.h>

#include "a3d_common.hpp"

namespace a3d {

namespace {

constexpr int kBlk = 256;

// Monotone float <-> uint32 map so atomicMin/atomicMax order floats.
__device__ __forceinline__ uint32_t f2ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Reduction slots (ordered-uint encoded): centroid lo[3] hi[3], vertex lo[3]
// hi[3], coord_max, pad_max.
enum { R_CLO = 0, R_CHI = 3, R_VLO = 6, R_VHI = 9, R_CMAX = 12, R_PAD = 13, R_WALK = 14, R_N = 16 };

struct Ws {
  uint32_t* red;
  uint64_t* keys_in;
  uint64_t* keys_out;
  uint32_t* vals_in;
  uint32_t* vals_out;
  float* leaf_box;   // [n][6] sorted order
  float* node_box;   // [n-1][6] leaf units and kept nodes
  int32_t* kl;       // [n-1] child refs: >=0 internal, <0 ~leafprim
  int32_t* kr;
  int32_t* first;    // [n-1]
  int32_t* size;     // [n-1]
  int32_t* par_int;  // [n-1]
  int32_t* par_leaf; // [n]
  int32_t* flag;     // [n-1]
  int32_t* kept;     // [n-1]
  int32_t* remap;    // [n-1]
  float* cost;       // [n-1] SAH cost of the subtree (units and kept nodes)
  int32_t* height;   // [n-1] kept nodes on the longest path down (units: 0)
  int32_t* front0;   // [n-1] frontier ping-pong
  int32_t* front1;
  int32_t* counts;   // [kRoundsMax + 1] frontier size per round
  int32_t* kl0;      // [n-1] fitted Karras tree, kept if a refinement is rejected
  int32_t* kr0;
  float* node_box0;  // [n-1][6]
  int32_t* leafy;    // [n-1] 1: the cost model makes this subtree one leaf
  void* cub;
  size_t cub_bytes;
  int64_t total;
};

constexpr int kRoundsMax = 4096;
// Deepest kept-node chain a walk stack is sized for (the Karras tree of
// distinct 95-bit keys is at most 96 deep).
constexpr int kMaxDepth = 96;
// Cost model (surface-area heuristic generalised to distance queries): a
// node is visited by a query whose search ball of radius r meets its box, so
// its weight is the volume of the box grown by r (Steiner:
// V + A r + pi W r^2 + 4/3 pi r^3, W = dx + dy + dz), divided by r; r -> 0
// gives the surface area A.  r = kWalkR x the scene diagonal.  kCi / kCt:
// cost of one internal-node visit (two child boxes) vs one triangle.
#ifndef A3D_WALK_R
#define A3D_WALK_R 0.03f
#endif
#ifndef A3D_CI
#define A3D_CI 1.2f
#endif
constexpr float kWalkR = A3D_WALK_R;
constexpr float kCi = A3D_CI, kCt = 1.0f;
constexpr int kTreeletLeaves = 7;

inline int64_t al(int64_t x) { return (x + 255) & ~int64_t(255); }

size_t cub_bytes_for(int64_t n) {
  size_t sort_b = 0, scan_b = 0;
  int nn = (int)(n > 1 ? n : 2);
  (void)rocprim::radix_sort_pairs(nullptr, sort_b, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                     (uint32_t*)nullptr, (uint32_t*)nullptr, nn, 0, 63);
  (void)a3d::exclusive_sum(nullptr, scan_b, (int32_t*)nullptr, (int32_t*)nullptr, nn);
  return sort_b > scan_b ? sort_b : scan_b;
}

Ws carve(void* base, int64_t n) {
  Ws w;
  int64_t m = n > 1 ? n - 1 : 1;
  int64_t off = 0;
  char* b = (char*)base;
  auto take = [&](int64_t bytes) { char* p = b ? b + off : nullptr; off += al(bytes); return p; };
  w.red = (uint32_t*)take(R_N * 4);
  w.keys_in = (uint64_t*)take(n * 8);
  w.keys_out = (uint64_t*)take(n * 8);
  w.vals_in = (uint32_t*)take(n * 4);
  w.vals_out = (uint32_t*)take(n * 4);
  w.leaf_box = (float*)take(n * 24);
  w.node_box = (float*)take(m * 24);
  w.kl = (int32_t*)take(m * 4);
  w.kr = (int32_t*)take(m * 4);
  w.first = (int32_t*)take(m * 4);
  w.size = (int32_t*)take(m * 4);
  w.par_int = (int32_t*)take(m * 4);
  w.par_leaf = (int32_t*)take(n * 4);
  w.flag = (int32_t*)take(m * 4);
  w.kept = (int32_t*)take(m * 4);
  w.remap = (int32_t*)take(m * 4);
  w.cost = (float*)take(m * 4);
  w.height = (int32_t*)take(m * 4);
  w.front0 = (int32_t*)take(m * 4);
  w.front1 = (int32_t*)take(m * 4);
  w.counts = (int32_t*)take((kRoundsMax + 1) * 4);
  w.kl0 = (int32_t*)take(m * 4);
  w.kr0 = (int32_t*)take(m * 4);
  w.node_box0 = (float*)take(m * 24);
  w.leafy = (int32_t*)take(m * 4);
  w.cub_bytes = cub_bytes_for(n);
  w.cub = take((int64_t)w.cub_bytes);
  w.total = off;
  return w;
}

int64_t blob_size(int64_t nF) {
  int64_t cap = nF > 1 ? nF - 1 : 1;
  return 256 + 64 * cap + 48 * nF + 32 * nF + 96 * cap + 48 * nF;
}

__device__ __forceinline__ v3 vtx(const float* __restrict__ V, int32_t i) {
  return mk(V[(int64_t)i * 3], V[(int64_t)i * 3 + 1], V[(int64_t)i * 3 + 2]);
}

// Conservative bound on how far the reference's plane-branch distance can fall
// below the true point-triangle distance: the computed normal
// cross(b-a, a-c) is off by ~3 ulp of |b-a||a-c| per component, i.e. by an
// angle of ~3*2^-24/sin(angle at a); over a triangle of diameter D that moves
// the plane by <= that angle * D.  16x margin.  Zero when the reference takes
// the degenerate (edge-only) branch.
__device__ __forceinline__ float tri_pad(v3 a, v3 b, v3 c) {
  v3 v21 = sub3(b, a), v13 = sub3(a, c), v32 = sub3(c, b);
  v3 nor = cross3(v21, v13);
  float nor_sq = dot3(nor, nor);
  if (nor_sq < 1e-12f) return 0.0f;
  float l21 = sqrtf(dot3(v21, v21)), l13 = sqrtf(dot3(v13, v13)), l32 = sqrtf(dot3(v32, v32));
  float s = sqrtf(nor_sq) / fmaxf(l21 * l13, 1e-30f);
  float diam = fmaxf(l21, fmaxf(l13, l32));
  return 16.0f * 5.9604645e-08f * diam * (1.0f + 1.0f / fmaxf(s, 1e-30f));
}

__global__ void k_init(uint32_t* red) {
  int t = threadIdx.x;
  if (t < R_N) {
    bool is_min = (t >= R_CLO && t < R_CHI) || (t >= R_VLO && t < R_VHI);
    red[t] = is_min ? 0xffffffffu : 0u;
  }
}

__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// Per face: centroid (reference formula, bvh_kernels.cu:898-902), bounds
// reductions: wave shuffles, then the block's waves through LDS, then one
// atomic per block per slot (the grid is capped, so a few thousand atomics
// in all).
__global__ __launch_bounds__(kBlk) void k_prep(const float* __restrict__ V, const int32_t* __restrict__ F,
                                                int64_t n, uint32_t* red) {
  __shared__ float s_r[kBlk / 64][R_N];
  float r[R_N];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    r[R_CLO + k] = 3.0e38f; r[R_CHI + k] = -3.0e38f; r[R_VLO + k] = 3.0e38f; r[R_VHI + k] = -3.0e38f;
  }
  r[R_CMAX] = 0.0f; r[R_PAD] = 0.0f;
  for (int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlk) {
    v3 a = vtx(V, F[i * 3]), b = vtx(V, F[i * 3 + 1]), c = vtx(V, F[i * 3 + 2]);
    float cx = (a.x + b.x + c.x) / 3.0f, cy = (a.y + b.y + c.y) / 3.0f, cz = (a.z + b.z + c.z) / 3.0f;
    r[R_CLO] = fminf(r[R_CLO], cx); r[R_CLO + 1] = fminf(r[R_CLO + 1], cy); r[R_CLO + 2] = fminf(r[R_CLO + 2], cz);
    r[R_CHI] = fmaxf(r[R_CHI], cx); r[R_CHI + 1] = fmaxf(r[R_CHI + 1], cy); r[R_CHI + 2] = fmaxf(r[R_CHI + 2], cz);
    r[R_VLO] = fminf(r[R_VLO], fminf(fminf(a.x, b.x), c.x));
    r[R_VLO + 1] = fminf(r[R_VLO + 1], fminf(fminf(a.y, b.y), c.y));
    r[R_VLO + 2] = fminf(r[R_VLO + 2], fminf(fminf(a.z, b.z), c.z));
    r[R_VHI] = fmaxf(r[R_VHI], fmaxf(fmaxf(a.x, b.x), c.x));
    r[R_VHI + 1] = fmaxf(r[R_VHI + 1], fmaxf(fmaxf(a.y, b.y), c.y));
    r[R_VHI + 2] = fmaxf(r[R_VHI + 2], fmaxf(fmaxf(a.z, b.z), c.z));
    r[R_PAD] = fmaxf(r[R_PAD], tri_pad(a, b, c));
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k <= R_PAD; ++k) {
    if (k == R_CMAX) continue;
    const bool is_min = (k >= R_CLO && k < R_CHI) || (k >= R_VLO && k < R_VHI);
    const float v = is_min ? wave_min(r[k]) : wave_max(r[k]);
    if (lane == 0) s_r[wv][k] = v;
  }
  __syncthreads();
  const int k = threadIdx.x;
  if (k <= R_PAD && k != R_CMAX) {
    const bool is_min = (k >= R_CLO && k < R_CHI) || (k >= R_VLO && k < R_VHI);
    float v = s_r[0][k];
    for (int w = 1; w < kBlk / 64; ++w) v = is_min ? fminf(v, s_r[w][k]) : fmaxf(v, s_r[w][k]);
    if (is_min) atomicMin(&red[k], f2ord(v));
    else atomicMax(&red[k], f2ord(v));
  }
}

__device__ __forceinline__ uint64_t spread21(uint64_t x) {
  x &= 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__device__ __forceinline__ uint32_t quant21(float c, float lo, float inv) {
  float t = (c - lo) * inv;
  t = fminf(fmaxf(t, 0.0f), 2097151.0f);
  return (uint32_t)t;
}

// 63-bit Morton key of each face centroid in the centroid bounding box.
__global__ __launch_bounds__(kBlk) void k_morton(const float* __restrict__ V, const int32_t* __restrict__ F,
                                                  int64_t n, const uint32_t* __restrict__ red,
                                                  uint64_t* keys, uint32_t* vals) {
  int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i >= n) return;
  float lo[3], inv[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = ord2f(red[R_CLO + k]);
    float ext = ord2f(red[R_CHI + k]) - lo[k];
    inv[k] = ext > 0.0f ? 2097151.0f / ext : 0.0f;
  }
  v3 a = vtx(V, F[i * 3]), b = vtx(V, F[i * 3 + 1]), c = vtx(V, F[i * 3 + 2]);
  float cx = (a.x + b.x + c.x) / 3.0f, cy = (a.y + b.y + c.y) / 3.0f, cz = (a.z + b.z + c.z) / 3.0f;
  uint64_t key = (spread21(quant21(cx, lo[0], inv[0])) << 2) | (spread21(quant21(cy, lo[1], inv[1])) << 1) |
                 spread21(quant21(cz, lo[2], inv[2]));
  keys[i] = key;
  vals[i] = (uint32_t)i;
}

// Common-prefix length of sorted keys i and j (Karras 2012 delta); equal keys
// are split by index, so every key is distinct.
__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  uint64_t a = k[i], b = k[j];
  if (a == b) return 64 + __clz((uint32_t)(i ^ j));
  return __clzll(a ^ b);
}

__global__ __launch_bounds__(kBlk) void k_karras(const uint64_t* __restrict__ keys, int64_t n, int32_t* kl,
                                                  int32_t* kr, int32_t* first, int32_t* size, int32_t* par_int,
                                                  int32_t* par_leaf, int32_t* flag) {
  int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i >= n - 1) return;
  int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(keys, n, i, i - d);
  int64_t lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int64_t j = i + l * d;
  int dnode = delta(keys, n, i, j);
  int64_t s = 0;
  int64_t t = l;
  do {
    t = (t + 1) / 2;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int64_t g = i + s * d + (d < 0 ? -1 : 0);
  int64_t lo = i < j ? i : j, hi = i < j ? j : i;
  int32_t left = (lo == g) ? ~(int32_t)g : (int32_t)g;
  int32_t right = (hi == g + 1) ? ~(int32_t)(g + 1) : (int32_t)(g + 1);
  kl[i] = left;
  kr[i] = right;
  first[i] = (int32_t)lo;
  size[i] = (int32_t)(hi - lo + 1);
  flag[i] = 0;
  if (left < 0) par_leaf[~left] = (int32_t)i; else par_int[left] = (int32_t)i;
  if (right < 0) par_leaf[~right] = (int32_t)i; else par_int[right] = (int32_t)i;
  if (i == 0) par_int[0] = -1;
}

__device__ __forceinline__ void leaf_box_of(const float* __restrict__ V, const int32_t* __restrict__ F,
                                            uint32_t face, float* o) {
  v3 a = vtx(V, F[(int64_t)face * 3]), b = vtx(V, F[(int64_t)face * 3 + 1]), c = vtx(V, F[(int64_t)face * 3 + 2]);
  float pad = tri_pad(a, b, c);
  o[0] = fminf(fminf(a.x, b.x), c.x) - pad;
  o[1] = fminf(fminf(a.y, b.y), c.y) - pad;
  o[2] = fminf(fminf(a.z, b.z), c.z) - pad;
  o[3] = fmaxf(fmaxf(a.x, b.x), c.x) + pad;
  o[4] = fmaxf(fmaxf(a.y, b.y), c.y) + pad;
  o[5] = fmaxf(fmaxf(a.z, b.z), c.z) + pad;
}

// Padded box of each triangle in sorted order.
__global__ __launch_bounds__(kBlk) void k_leaf_boxes(const float* __restrict__ V, const int32_t* __restrict__ F,
                                                      const uint32_t* __restrict__ order, int64_t n,
                                                      float* leaf_box) {
  int64_t j = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (j >= n) return;
  float b[6];
  leaf_box_of(V, F, order[j], b);
  for (int k = 0; k < 6; ++k) leaf_box[j * 6 + k] = b[k];
}

// Search radius of the cost model, from the vertex bounds (after k_prep).
__global__ void k_walk_radius(uint32_t* red) {
  if (threadIdx.x != 0) return;
  float d2 = 0.0f;
  for (int k = 0; k < 3; ++k) {
    const float e = ord2f(red[R_VHI + k]) - ord2f(red[R_VLO + k]);
    d2 += e * e;
  }
  red[R_WALK] = __float_as_uint(kWalkR * sqrtf(d2));
}

__device__ __forceinline__ float measure(const float* b, float r) {
  const float dx = fmaxf(b[3] - b[0], 0.0f), dy = fmaxf(b[4] - b[1], 0.0f), dz = fmaxf(b[5] - b[2], 0.0f);
  const float a = 2.0f * (dx * dy + dy * dz + dz * dx);
  if (!(r > 0.0f)) return a;
  return dx * dy * dz / r + a + 3.14159265f * (dx + dy + dz) * r + 4.18879020f * r * r;
}

__device__ __forceinline__ bool ref_kept(int32_t ref, const int32_t* __restrict__ kept) {
  return ref >= 0 && kept[ref] != 0;
}
__device__ __forceinline__ const float* ref_box(int32_t ref, const float* leaf_box, const float* node_box) {
  return ref < 0 ? leaf_box + (int64_t)(~ref) * 6 : node_box + (int64_t)ref * 6;
}
__device__ __forceinline__ float ref_cost(int32_t ref, const float* leaf_box, const float* cost, float r) {
  return ref < 0 ? kCt * measure(leaf_box + (int64_t)(~ref) * 6, r) : cost[ref];
}
__device__ __forceinline__ int32_t ref_height(int32_t ref, const int32_t* height) {
  return ref < 0 ? 0 : height[ref];
}

// Kept (emitted) internal nodes: not made a leaf by the cost model and with
// no ancestor made one.  Only nodes of <= leaf_max triangles can be leaves,
// so the climb stops at the first larger ancestor.
__global__ __launch_bounds__(kBlk) void k_kept(int64_t m, int leaf_max, const int32_t* __restrict__ size,
                                                const int32_t* __restrict__ leafy,
                                                const int32_t* __restrict__ par_int, int32_t* kept) {
  int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i >= m) return;
  int k = leafy[i] ? 0 : 1;
  for (int32_t p = par_int[i]; k && p >= 0 && size[p] <= leaf_max; p = par_int[p])
    if (leafy[p]) k = 0;
  kept[i] = k;
}

__global__ void k_zero_i32(int32_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}

// Start of a bottom-up pass over the kept nodes: flag = number of children
// already final (units and triangles); nodes with two such children form the
// first frontier.  counts[] must be zero.
// Appends `val` for every thread of the block whose `pred` holds: one
// device-scope atomic per block (atomics on one address serialise at the
// memory side, so per-wave atomics cost ~10 ns x the wave count).  Every
// thread of the block must call it.
__device__ __forceinline__ void block_append(bool pred, int32_t val, int32_t* count, int32_t* out) {
  __shared__ int32_t s_cnt[kBlk / 64];
  __shared__ int32_t s_base;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t m = __ballot(pred);
  if (lane == 0) s_cnt[wv] = (int32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t t = 0;
    for (int w = 0; w < kBlk / 64; ++w) { const int32_t c = s_cnt[w]; s_cnt[w] = t; t += c; }
    s_base = t > 0 ? atomicAdd(count, t) : 0;
  }
  __syncthreads();
  if (pred)
    out[s_base + s_cnt[wv] +
        (int32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = val;
  __syncthreads();  // s_cnt / s_base are reused by the next call
}

// Grid-stride loops whose trip count is uniform over each block (for
// block_append): the block's threads cover [i0, i0 + kBlk).
#define A3D_BLOCK_LOOP(i0, n) \
  for (int64_t i0 = (int64_t)blockIdx.x * kBlk; i0 < (n); i0 += (int64_t)gridDim.x * kBlk)

__global__ __launch_bounds__(kBlk) void k_pass_init(int64_t m, const int32_t* __restrict__ kept,
                                                     const int32_t* __restrict__ kl, const int32_t* __restrict__ kr,
                                                     int32_t* flag, int32_t* counts, int32_t* front) {
  A3D_BLOCK_LOOP(i0, m) {
    const int64_t i = i0 + threadIdx.x;
    bool ready = false;
    if (i < m && kept[i]) {
      const int c = (ref_kept(kl[i], kept) ? 0 : 1) + (ref_kept(kr[i], kept) ? 0 : 1);
      flag[i] = c;
      ready = c == 2;
    }
    block_append(ready, (int32_t)i, &counts[0], front);
  }
}

// Start of the bottom-up pass over every internal node: triangles are final.
// Its first round takes the nodes with two triangle children straight from
// the node array (no list).
__global__ __launch_bounds__(kBlk) void k_pass_init_full(int64_t m, const int32_t* __restrict__ kl,
                                                          const int32_t* __restrict__ kr, int32_t* flag) {
  const int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i < m) flag[i] = (kl[i] < 0 ? 1 : 0) + (kr[i] < 0 ? 1 : 0);
}

// The second child to finish puts the parent into the next round's frontier
// (single-lane form, for the treelet rounds).
__device__ __forceinline__ void arrive(int32_t node, int r, const int32_t* __restrict__ par_int, int32_t* flag,
                                       int32_t* counts, int32_t* fout) {
  const int32_t p = par_int[node];
  if (p < 0) return;
  if (atomicAdd(&flag[p], 1) == 1) fout[atomicAdd(&counts[r + 1], 1)] = p;
}

// Box, cost and height of `node` from its children.  FULL (first pass over
// every internal node): a node of <= leaf_max triangles becomes a leaf when
// that is cheaper (leafy = 1; its cost is the leaf cost, its height 0).
template <bool FULL>
__device__ __forceinline__ void fit_node(int32_t node, const int32_t* __restrict__ kl,
                                         const int32_t* __restrict__ kr, const int32_t* __restrict__ size,
                                         int leaf_max, float r, const float* leaf_box, float* node_box,
                                         float* cost, int32_t* height, int32_t* leafy) {
  const int32_t a = kl[node], b = kr[node];
  const float* ba = ref_box(a, leaf_box, node_box);
  const float* bb = ref_box(b, leaf_box, node_box);
  float o[6];
  for (int k = 0; k < 3; ++k) { o[k] = fminf(ba[k], bb[k]); o[k + 3] = fmaxf(ba[k + 3], bb[k + 3]); }
  for (int k = 0; k < 6; ++k) node_box[(int64_t)node * 6 + k] = o[k];
  const float mu = measure(o, r);
  const float split = kCi * mu + ref_cost(a, leaf_box, cost, r) + ref_cost(b, leaf_box, cost, r);
  const int32_t ha = ref_height(a, height), hb = ref_height(b, height);
  if (FULL) {
    const int32_t nt = size[node];
    const bool leaf = nt <= leaf_max && kCt * (float)nt * mu <= split;
    leafy[node] = leaf ? 1 : 0;
    cost[node] = leaf ? kCt * (float)nt * mu : split;
    height[node] = leaf ? 0 : 1 + (ha > hb ? ha : hb);
  } else {
    cost[node] = split;
    height[node] = 1 + (ha > hb ? ha : hb);
  }
}

// One fitting round over every internal node: one thread per frontier node.
__global__ __launch_bounds__(kBlk) void k_fit_round(int r, int64_t m, int32_t* counts, const int32_t* __restrict__ fin,
                                                     int32_t* fout, const int32_t* __restrict__ kl,
                                                     const int32_t* __restrict__ kr,
                                                     const int32_t* __restrict__ par_int,
                                                     const int32_t* __restrict__ size, int leaf_max,
                                                     const uint32_t* __restrict__ red,
                                                     const float* __restrict__ leaf_box, float* node_box,
                                                     float* cost, int32_t* height, int32_t* flag, int32_t* leafy) {
  // round 0: every node with two triangle children, read from the node array
  const int64_t nf = r == 0 ? m : counts[r];
  const float wr = __uint_as_float(red[R_WALK]);
  A3D_BLOCK_LOOP(i0, nf) {
    const int64_t i = i0 + threadIdx.x;
    int32_t p = -1;
    if (i < nf && (r > 0 || (kl[i] < 0 && kr[i] < 0))) {
      const int32_t node = r == 0 ? (int32_t)i : fin[i];
      fit_node<true>(node, kl, kr, size, leaf_max, wr, leaf_box, node_box, cost, height, leafy);
      p = par_int[node];
      if (p >= 0 && atomicAdd(&flag[p], 1) != 1) p = -1;
    }
    block_append(p >= 0, p, &counts[r + 1], fout);
  }
}

// ------------------------------------------------------------ treelets
// A treelet of node n: start from n's two children and repeatedly replace
// the kept child of largest surface area by its two children, up to 7
// leaves.  The wave finds the SAH-optimal binary tree over those leaves by
// dynamic programming over the 127 leaf subsets (subsets of one size in
// parallel, sizes in order) and, when it is cheaper, rebuilds the treelet in
// place on the same internal nodes.
struct TreeletLds {
  float area[128];
  float cost[128];
  int32_t part[128];
  int32_t hsub[128];
  float lbox[8][6];
  float larea[8];
  float lcost[8];
  int32_t lheight[8];
  int32_t leaves[8];
  int32_t inner[8];
  int32_t st_set[8];
  int32_t st_slot[8];
  int32_t nl;
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void set_box(const TreeletLds& T, int32_t S, float* o) {
  o[0] = o[1] = o[2] = 3.0e38f;
  o[3] = o[4] = o[5] = -3.0e38f;
  for (int j = 0; j < kTreeletLeaves; ++j) {
    if (!((S >> j) & 1)) continue;
    for (int k = 0; k < 3; ++k) { o[k] = fminf(o[k], T.lbox[j][k]); o[k + 3] = fmaxf(o[k + 3], T.lbox[j][k + 3]); }
  }
}

// Returns true when the treelet was rebuilt (node's box, cost and height are
// then written too).  Every lane of the wave calls it with the same node.
__device__ bool treelet(int32_t node, TreeletLds& T, int lane, float wr, int32_t* kl, int32_t* kr,
                        int32_t* par_int, int32_t* par_leaf, const int32_t* __restrict__ kept,
                        const float* __restrict__ leaf_box, float* node_box, float* cost, int32_t* height) {
  if (lane == 0) {
    int nl = 2, ni = 1;
    T.inner[0] = node;
    T.leaves[0] = kl[node];
    T.leaves[1] = kr[node];
    for (int j = 0; j < 2; ++j) {
      const float* b = ref_box(T.leaves[j], leaf_box, node_box);
      for (int k = 0; k < 6; ++k) T.lbox[j][k] = b[k];
      T.larea[j] = measure(b, wr);
    }
    while (nl < kTreeletLeaves) {
      int best = -1;
      float ba = -1.0f;
      for (int j = 0; j < nl; ++j)
        if (ref_kept(T.leaves[j], kept) && T.larea[j] > ba) { ba = T.larea[j]; best = j; }
      if (best < 0) break;
      const int32_t c = T.leaves[best];
      T.inner[ni++] = c;
      T.leaves[best] = kl[c];
      T.leaves[nl] = kr[c];
      const int js[2] = {best, nl};
      for (int q = 0; q < 2; ++q) {
        const float* b = ref_box(T.leaves[js[q]], leaf_box, node_box);
        for (int k = 0; k < 6; ++k) T.lbox[js[q]][k] = b[k];
        T.larea[js[q]] = measure(b, wr);
      }
      ++nl;
    }
    for (int j = 0; j < nl; ++j) {
      T.lcost[j] = ref_cost(T.leaves[j], leaf_box, cost, wr);
      T.lheight[j] = ref_height(T.leaves[j], height);
    }
    T.nl = nl;
  }
  wave_sync();
  const int nl = T.nl;
  if (nl < 3) return false;
  const int nsub = 1 << nl;
  for (int S = lane; S < nsub; S += 64) {
    if (S == 0) continue;
    float b[6];
    set_box(T, S, b);
    T.area[S] = measure(b, wr);
    if (__popc(S) == 1) {
      const int j = __ffs(S) - 1;
      T.cost[S] = T.lcost[j];
      T.hsub[S] = T.lheight[j];
      T.part[S] = 0;
    }
  }
  wave_sync();
  for (int sz = 2; sz <= nl; ++sz) {
    for (int S = lane; S < nsub; S += 64) {
      if (__popc(S) != sz) continue;
      const int low = S & -S;
      float best = 3.0e38f;
      int bp = 0;
      for (int P = (S - 1) & S; P; P = (P - 1) & S) {
        if (!(P & low)) continue;
        const float c = T.cost[P] + T.cost[S ^ P];
        if (c < best) { best = c; bp = P; }
      }
      T.cost[S] = kCi * T.area[S] + best;
      T.part[S] = bp;
      const int32_t h0 = T.hsub[bp], h1 = T.hsub[S ^ bp];
      T.hsub[S] = 1 + (h0 > h1 ? h0 : h1);
    }
    wave_sync();
  }
  const int full = nsub - 1;
  const float cur = kCi * T.area[full] + ref_cost(kl[node], leaf_box, cost, wr) + ref_cost(kr[node], leaf_box, cost, wr);
  // Non-finite costs (huge or NaN coordinates) leave partitions unset: keep
  // the treelet as it is.
  if (!(T.cost[full] < cur * (1.0f - 1e-5f)) || !(cur < 3.0e38f)) return false;
  if (lane == 0) {
    int sp = 1, next = 1;
    T.st_set[0] = full;
    T.st_slot[0] = node;
    while (sp > 0) {
      --sp;
      const int S = T.st_set[sp];
      const int32_t slot = T.st_slot[sp];
      const int sub[2] = {T.part[S], S ^ T.part[S]};
      int32_t ref[2];
      for (int q = 0; q < 2; ++q) {
        if (__popc(sub[q]) == 1) {
          ref[q] = T.leaves[__ffs(sub[q]) - 1];
        } else {
          ref[q] = T.inner[next++];
          T.st_set[sp] = sub[q];
          T.st_slot[sp] = ref[q];
          ++sp;
        }
        if (ref[q] < 0) par_leaf[~ref[q]] = slot;
        else par_int[ref[q]] = slot;
      }
      kl[slot] = ref[0];
      kr[slot] = ref[1];
      float b[6];
      set_box(T, S, b);
      for (int k = 0; k < 6; ++k) node_box[(int64_t)slot * 6 + k] = b[k];
      cost[slot] = T.cost[S];
      height[slot] = T.hsub[S];
    }
  }
  return true;
}

// One restructuring round: one wavefront per frontier node.
__global__ __launch_bounds__(kBlk) void k_treelet_round(int r, int32_t* counts, const int32_t* __restrict__ fin,
                                                         int32_t* fout, int32_t* kl, int32_t* kr, int32_t* par_int,
                                                         int32_t* par_leaf, const int32_t* __restrict__ kept,
                                                         const uint32_t* __restrict__ red,
                                                         const float* __restrict__ leaf_box, float* node_box,
                                                         float* cost, int32_t* height, int32_t* flag) {
  __shared__ TreeletLds s_t[kBlk / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int32_t nf = counts[r];
  const float wr = __uint_as_float(red[R_WALK]);
  const int64_t nw = (int64_t)gridDim.x * (kBlk / 64);
  for (int64_t i = blockIdx.x * (int64_t)(kBlk / 64) + wv; i < nf; i += nw) {
    const int32_t node = fin[i];
    const bool rebuilt = treelet(node, s_t[wv], lane, wr, kl, kr, par_int, par_leaf, kept, leaf_box, node_box,
                                 cost, height);
    if (lane == 0) {
      if (!rebuilt) fit_node<false>(node, kl, kr, nullptr, 0, wr, leaf_box, node_box, cost, height, nullptr);
      arrive(node, r, par_int, flag, counts, fout);
    }
    wave_sync();
  }
}

__device__ __forceinline__ void child_info(int32_t c, const int32_t* __restrict__ kept,
                                           const int32_t* __restrict__ size,
                                           const int32_t* __restrict__ first, const int32_t* __restrict__ remap,
                                           const float* __restrict__ leaf_box, const float* __restrict__ node_box,
                                           int32_t& ref, const float*& box) {
  if (c < 0) {
    ref = leaf_ref(~c, 1);
    box = leaf_box + (int64_t)(~c) * 6;
  } else if (!kept[c]) {
    ref = leaf_ref(first[c], size[c]);
    box = node_box + (int64_t)c * 6;
  } else {
    ref = remap[c];
    box = node_box + (int64_t)c * 6;
  }
}

__global__ __launch_bounds__(kBlk) void k_emit_nodes(int64_t m, const int32_t* __restrict__ kept,
                                                      const int32_t* __restrict__ remap,
                                                      const int32_t* __restrict__ kl, const int32_t* __restrict__ kr,
                                                      const int32_t* __restrict__ size,
                                                      const int32_t* __restrict__ first,
                                                      const float* __restrict__ leaf_box,
                                                      const float* __restrict__ node_box, Node64* out) {
  int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i >= m || !kept[i]) return;
  int32_t lr, rr;
  const float *lb, *rb;
  child_info(kl[i], kept, size, first, remap, leaf_box, node_box, lr, lb);
  child_info(kr[i], kept, size, first, remap, leaf_box, node_box, rr, rb);
  Node64 nd;
  nd.a = make_float4(lb[0], lb[1], lb[2], lb[3]);
  nd.b = make_float4(lb[4], lb[5], rb[0], rb[1]);
  nd.c = make_float4(rb[2], rb[3], rb[4], rb[5]);
  nd.d = make_int4(lr, rr, 0, 0);
  out[remap[i]] = nd;
}

__global__ __launch_bounds__(kBlk) void k_emit_tris(const float* __restrict__ V, const int32_t* __restrict__ F,
                                                     const uint32_t* __restrict__ order, int64_t n, Tri48* out,
                                                     TriBox32* boxes) {
  int64_t j = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (j >= n) return;
  uint32_t f = order[j];
  v3 a = vtx(V, F[(int64_t)f * 3]), b = vtx(V, F[(int64_t)f * 3 + 1]), c = vtx(V, F[(int64_t)f * 3 + 2]);
  Tri48 t;
  t.p = make_float4(a.x, a.y, a.z, b.x);
  t.q = make_float4(b.y, b.z, c.x, c.y);
  const float pad = tri_pad(a, b, c);
  t.r = make_float4(c.z, __int_as_float((int32_t)f), pad, 0.0f);
  out[j] = t;
  boxes[j] = tri_box_of(a, b, c, pad);
}

// Meshes with <= leaf_max faces (including none): one root whose left child
// is a leaf holding every face and whose right child is an empty box.
__global__ void k_small_root(const float* __restrict__ V, const int32_t* __restrict__ F, int64_t n, Node64* out,
                             Tri48* tris, TriBox32* boxes) {
  if (threadIdx.x != 0) return;
  float b[6] = {3.0e38f, 3.0e38f, 3.0e38f, -3.0e38f, -3.0e38f, -3.0e38f};
  for (int64_t j = 0; j < n; ++j) {
    float q[6];
    leaf_box_of(V, F, (uint32_t)j, q);
    for (int k = 0; k < 3; ++k) { b[k] = fminf(b[k], q[k]); b[k + 3] = fmaxf(b[k + 3], q[k + 3]); }
    v3 a = vtx(V, F[j * 3]), bb = vtx(V, F[j * 3 + 1]), c = vtx(V, F[j * 3 + 2]);
    Tri48 t;
    t.p = make_float4(a.x, a.y, a.z, bb.x);
    t.q = make_float4(bb.y, bb.z, c.x, c.y);
    const float pad = tri_pad(a, bb, c);
    t.r = make_float4(c.z, __int_as_float((int32_t)j), pad, 0.0f);
    tris[j] = t;
    boxes[j] = tri_box_of(a, bb, c, pad);
  }
  const float inf = __int_as_float(0x7f800000);
  Node64 nd;
  if (n == 0) { b[0] = b[1] = b[2] = inf; b[3] = b[4] = b[5] = -inf; }
  nd.a = make_float4(b[0], b[1], b[2], b[3]);
  nd.b = make_float4(b[4], b[5], inf, inf);
  nd.c = make_float4(inf, -inf, -inf, -inf);
  int32_t cnt = n > 0 ? (int32_t)n : 1;
  nd.d = make_int4(leaf_ref(0, cnt), leaf_ref(0, 1), 0, 0);
  out[0] = nd;
}

// Oriented box of one child (see Obb96): its triangles are the leaves of the
// subtree under `ref`, gathered by a DFS over node refs (no more than
// kObbCap triangles, else no box).  n = the normalised sum of the triangles'
// area vectors (FP64), t1 = n x (the world axis least aligned with n),
// normalised; both shrunk by 2^-21 so they stay within unit length after
// rounding to float; t2 = obb_t2(n, t1) in float exactly as the walks compute
// it.  The ranges of a . v over the vertices (FP64) are widened by the
// largest triangle pad and rounded outward to float.
__device__ void child_obb(const Node64* __restrict__ nodes, const Tri48* __restrict__ tris, int32_t ref,
                          float4* out) {
  const float inf = __int_as_float(0x7f800000);
  out[0] = make_float4(0.0f, 0.0f, 0.0f, -inf);
  out[1] = make_float4(0.0f, 0.0f, 0.0f, -inf);
  out[2] = make_float4(inf, inf, -inf, inf);
  int32_t stk[64];
  double sx = 0.0, sy = 0.0, sz = 0.0;
  double ax[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, pad = 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1) {
      const double len = sqrt(sx * sx + sy * sy + sz * sz);
      if (!(len > 0.0)) return;
      const double k = 1.0 - 4.76837158203125e-07;  // 1 - 2^-21
      const double n[3] = {sx / len, sy / len, sz / len};
      const int e = (fabs(n[0]) <= fabs(n[1]) && fabs(n[0]) <= fabs(n[2])) ? 0 : (fabs(n[1]) <= fabs(n[2]) ? 1 : 2);
      double t[3] = {0.0, 0.0, 0.0};  // n x e_axis
      t[(e + 1) % 3] = n[(e + 2) % 3];
      t[(e + 2) % 3] = -n[(e + 1) % 3];
      const double tl = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
      const v3 nf = mk((float)(n[0] * k), (float)(n[1] * k), (float)(n[2] * k));
      const v3 tf = mk((float)(t[0] / tl * k), (float)(t[1] / tl * k), (float)(t[2] / tl * k));
      const v3 uf = obb_t2(nf, tf);
      const v3 a3[3] = {nf, tf, uf};
      for (int q = 0; q < 3; ++q) {
        ax[q][0] = a3[q].x;
        ax[q][1] = a3[q].y;
        ax[q][2] = a3[q].z;
      }
      out[0] = make_float4(nf.x, nf.y, nf.z, -inf);
      out[1] = make_float4(tf.x, tf.y, tf.z, -inf);
    }
    int sp = 0, total = 0;
    int32_t cur = ref;
    while (true) {
      if (cur >= 0) {
        if (sp >= 64) return;
        stk[sp++] = nodes[cur].d.y;
        cur = nodes[cur].d.x;
        continue;
      }
      const int32_t s0 = leaf_start(cur), cnt = leaf_count(cur);
      total += cnt;
      if (total > kObbCap) return;
      for (int32_t t = s0; t < s0 + cnt; ++t) {
        v3 a, b, c;
        int32_t id;
        load_tri(tris, t, a, b, c, id);
        if (pass == 0) {
          const double ux = (double)b.x - a.x, uy = (double)b.y - a.y, uz = (double)b.z - a.z;
          const double vx = (double)c.x - a.x, vy = (double)c.y - a.y, vz = (double)c.z - a.z;
          sx += uy * vz - uz * vy;
          sy += uz * vx - ux * vz;
          sz += ux * vy - uy * vx;
        } else {
          const v3 vv[3] = {a, b, c};
          for (int q = 0; q < 3; ++q)
            for (int r = 0; r < 3; ++r) {
              const double d = ax[r][0] * vv[q].x + ax[r][1] * vv[q].y + ax[r][2] * vv[q].z;
              lo[r] = d < lo[r] ? d : lo[r];
              hi[r] = d > hi[r] ? d : hi[r];
            }
          const double pd = tri_pad_of(tris, t);
          pad = pd > pad ? pd : pad;
        }
      }
      if (sp == 0) break;
      cur = stk[--sp];
    }
  }
  float lf[3], hf[3];
  for (int r = 0; r < 3; ++r) {
    const double m = 1e-12 * (fabs(lo[r]) + fabs(hi[r])) + pad;
    const double l = lo[r] - m, h = hi[r] + m;
    lf[r] = (float)l;
    hf[r] = (float)h;
    if ((double)lf[r] > l) lf[r] = nextafterf(lf[r], -inf);
    if ((double)hf[r] < h) hf[r] = nextafterf(hf[r], inf);
  }
  out[0].w = lf[0];
  out[1].w = lf[1];
  out[2] = make_float4(hf[0], hf[1], lf[2], hf[2]);
}

// Oriented box of one triangle (see tobbs_of): FP64 frame from the
// vertices, float axes shrunk by 2^-21, FP64 ranges widened by `pad` and
// rounded outward.
__device__ void tri_obb(v3 a, v3 b, v3 c, float pad, float4* out) {
  const float inf = __int_as_float(0x7f800000);
  out[0] = make_float4(0.0f, 0.0f, 0.0f, -inf);
  out[1] = make_float4(0.0f, 0.0f, 0.0f, -inf);
  out[2] = make_float4(inf, inf, -inf, inf);
  const double ux = (double)b.x - a.x, uy = (double)b.y - a.y, uz = (double)b.z - a.z;
  const double vx = (double)c.x - a.x, vy = (double)c.y - a.y, vz = (double)c.z - a.z;
  const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
  const double nl = sqrt(nx * nx + ny * ny + nz * nz), ul = sqrt(ux * ux + uy * uy + uz * uz);
  if (!(nl > 0.0) || !(ul > 0.0)) return;
  const double k = 1.0 - 4.76837158203125e-07;  // 1 - 2^-21
  const v3 nf = mk((float)(nx / nl * k), (float)(ny / nl * k), (float)(nz / nl * k));
  const v3 tf = mk((float)(ux / ul * k), (float)(uy / ul * k), (float)(uz / ul * k));
  const v3 af[3] = {nf, tf, obb_t2(nf, tf)};
  const v3 vv[3] = {a, b, c};
  float lf[3], hf[3];
  for (int r = 0; r < 3; ++r) {
    double lo = 1e300, hi = -1e300;
    for (int q = 0; q < 3; ++q) {
      const double d = (double)af[r].x * vv[q].x + (double)af[r].y * vv[q].y + (double)af[r].z * vv[q].z;
      lo = d < lo ? d : lo;
      hi = d > hi ? d : hi;
    }
    const double m = 1e-12 * (fabs(lo) + fabs(hi)) + (double)pad;
    lo -= m;
    hi += m;
    lf[r] = (float)lo;
    hf[r] = (float)hi;
    if ((double)lf[r] > lo) lf[r] = nextafterf(lf[r], -inf);
    if ((double)hf[r] < hi) hf[r] = nextafterf(hf[r], inf);
  }
  out[0] = make_float4(nf.x, nf.y, nf.z, lf[0]);
  out[1] = make_float4(tf.x, tf.y, tf.z, lf[1]);
  out[2] = make_float4(hf[0], hf[1], lf[2], hf[2]);
}

__global__ __launch_bounds__(kBlk) void k_tri_obbs(void* blob) {
  const BvhHeader* H = header(blob);
  const int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= (int64_t)H->n_tris) return;
  v3 a, b, c;
  int32_t id;
  const Tri48* tris = tris_of(blob);
  load_tri(tris, (int32_t)t, a, b, c, id);
  float4* out = reinterpret_cast<float4*>((char*)blob + H->tobbs_off) + 3 * t;
  float4 o[3];
  tri_obb(a, b, c, tri_pad_of(tris, (int32_t)t), o);
  out[0] = o[0];
  out[1] = o[1];
  out[2] = o[2];
}

// One thread per (node, child) of the emitted tree.
__global__ __launch_bounds__(kBlk) void k_obbs(void* blob) {
  const BvhHeader* H = header(blob);
  const int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= 2 * (int64_t)H->n_nodes) return;
  const int64_t node = t >> 1;
  const int side = (int)(t & 1);
  const Node64* nodes = nodes_of(blob);
  Obb96* ob = reinterpret_cast<Obb96*>((char*)blob + H->obbs_off);
  float4 c[3];
  if (H->n_tris > 0) {
    child_obb(nodes, tris_of(blob), side ? nodes[node].d.y : nodes[node].d.x, c);
  } else {
    const float inf = __int_as_float(0x7f800000);
    c[0] = make_float4(0.0f, 0.0f, 0.0f, -inf);
    c[1] = make_float4(0.0f, 0.0f, 0.0f, -inf);
    c[2] = make_float4(inf, inf, -inf, inf);
  }
  for (int q = 0; q < 3; ++q) ob[node].c[3 * side + q] = c[q];
}

// depth: the root's height (kept nodes on the longest path down), or null
// for the single-leaf tree.
__global__ void k_header(BvhHeader* h, int64_t n, int64_t cap, int leaf_max, const uint32_t* __restrict__ red,
                         const int32_t* __restrict__ kept, const int32_t* __restrict__ remap, int64_t m,
                         const int32_t* __restrict__ depth) {
  if (threadIdx.x != 0) return;
  BvhHeader H;
  for (int k = 0; k < 40; ++k) H.pad[k] = 0;
  H.magic = kBvhMagic;
  H.n_tris = (int32_t)n;
  H.n_nodes = (n > leaf_max && m > 0) ? remap[m - 1] + kept[m - 1] : 1;
  H.n_node_cap = (int32_t)cap;
  H.nodes_off = 256;
  H.tris_off = 256 + 64 * cap;
  H.boxes_off = H.tris_off + 48 * n;
  H.obbs_off = H.boxes_off + 32 * n;
  H.tobbs_off = H.obbs_off + 96 * cap;
  float cm = 0.0f;
  for (int k = 0; k < 3; ++k) {
    H.scene_lo[k] = n > 0 ? ord2f(red[R_VLO + k]) : 0.0f;
    H.scene_hi[k] = n > 0 ? ord2f(red[R_VHI + k]) : 0.0f;
    cm = fmaxf(cm, fmaxf(fabsf(H.scene_lo[k]), fabsf(H.scene_hi[k])));
  }
  H.coord_max = cm;
  H.tri_pad_max = n > 0 ? ord2f(red[R_PAD]) : 0.0f;
  H.leaf_max = leaf_max;
  H.depth = depth ? depth[0] : 1;
  *h = H;
}

}  // namespace

}  // namespace a3d

using namespace a3d;

extern "C" int a3d_bvh_sizes(int64_t nF, int leaf_max, int64_t* blob_bytes, int64_t* ws_bytes) {
  if (nF < 0 || nF > (int64_t(1) << 28)) return fail(A3D_ERR_ARG, "a3d_bvh_sizes: nF=%lld out of range [0, 2^28]", (long long)nF);
  if (leaf_max < 1 || leaf_max > 8) return fail(A3D_ERR_ARG, "a3d_bvh_sizes: leaf_max=%d not in [1, 8]", leaf_max);
  if (!blob_bytes || !ws_bytes) return fail(A3D_ERR_ARG, "a3d_bvh_sizes: null output");
  *blob_bytes = blob_size(nF);
  *ws_bytes = carve(nullptr, nF > 0 ? nF : 1).total;
  return A3D_OK;
}

namespace {

// Launches `count` frontier rounds starting at round r0: TREELET false fits
// every internal node (and chooses the leaves), true restructures the kept
// nodes in treelets.
template <bool TREELET>
int launch_rounds(const Ws& w, int64_t m, int r0, int count, int leaf_max, hipStream_t s) {
  const unsigned g_fit = 1024, g_tree = 2048;
  for (int r = r0; r < r0 + count; ++r) {
    int32_t* fin = (r & 1) ? w.front1 : w.front0;
    int32_t* fout = (r & 1) ? w.front0 : w.front1;
    if (TREELET)
      hipLaunchKernelGGL(k_treelet_round, dim3(g_tree), dim3(kBlk), 0, s, r, w.counts, fin, fout, w.kl, w.kr,
                         w.par_int, w.par_leaf, w.kept, (const uint32_t*)w.red, w.leaf_box, w.node_box, w.cost,
                         w.height, w.flag);
    else
      hipLaunchKernelGGL(k_fit_round, dim3(g_fit), dim3(kBlk), 0, s, r, m, w.counts, fin, fout, w.kl, w.kr,
                         w.par_int, w.size, leaf_max, (const uint32_t*)w.red, w.leaf_box, w.node_box, w.cost,
                         w.height, w.flag, w.leafy);
    A3D_LAUNCH_CHECK();
  }
  return A3D_OK;
}

int pass_init(const Ws& w, int64_t m, bool full, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_i32, dim3(16), dim3(kBlk), 0, s, w.counts, (int64_t)(kRoundsMax + 1));
  A3D_LAUNCH_CHECK();
  const unsigned g = (unsigned)(div_up(m, kBlk) < 1024 ? div_up(m, kBlk) : 1024);
  if (full)
    hipLaunchKernelGGL(k_pass_init_full, dim3((unsigned)div_up(m, kBlk)), dim3(kBlk), 0, s, m, w.kl, w.kr, w.flag);
  else
    hipLaunchKernelGGL(k_pass_init, dim3(g), dim3(kBlk), 0, s, m, w.kept, w.kl, w.kr, w.flag, w.counts, w.front0);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

}  // namespace

extern "C" int a3d_bvh_build_ex(const float* V, int64_t nV, const int32_t* F, int64_t nF, int leaf_max,
                                int treelet_passes, void* blob, int64_t blob_bytes, void* ws, int64_t ws_bytes,
                                a3d_stream_t stream) {
  hipStream_t s = (hipStream_t)stream;
  int64_t need_blob = 0, need_ws = 0;
  int rc = a3d_bvh_sizes(nF, leaf_max, &need_blob, &need_ws);
  if (rc) return rc;
  if (treelet_passes < 0 || treelet_passes > 8)
    return fail(A3D_ERR_ARG, "a3d_bvh_build: treelet_passes=%d not in [0, 8]", treelet_passes);
  if (!blob || !ws) return fail(A3D_ERR_ARG, "a3d_bvh_build: null blob/workspace");
  if ((nF > 0) && (!V || !F || nV <= 0)) return fail(A3D_ERR_ARG, "a3d_bvh_build: null or empty mesh arrays");
  if (blob_bytes < need_blob) return fail(A3D_ERR_CAPACITY, "a3d_bvh_build: blob %lld < %lld bytes", (long long)blob_bytes, (long long)need_blob);
  if (ws_bytes < need_ws) return fail(A3D_ERR_CAPACITY, "a3d_bvh_build: workspace %lld < %lld bytes", (long long)ws_bytes, (long long)need_ws);
  if (((uintptr_t)blob & 255) || ((uintptr_t)ws & 255)) return fail(A3D_ERR_ARG, "a3d_bvh_build: blob/workspace must be 256-byte aligned");
  int64_t n = nF;
  int64_t cap = n > 1 ? n - 1 : 1;
  Ws w = carve(ws, n > 0 ? n : 1);
  char* base = (char*)blob;
  Node64* nodes = (Node64*)(base + 256);
  Tri48* tris = (Tri48*)(base + 256 + 64 * cap);
  TriBox32* boxes = (TriBox32*)(base + 256 + 64 * cap + 48 * n);
  int64_t m = n - 1;

  hipLaunchKernelGGL(k_init, dim3(1), dim3(64), 0, s, w.red);
  A3D_LAUNCH_CHECK();
  if (n > 0) {
    int64_t g = div_up(n, kBlk);
    if (g > 1024) g = 1024;
    hipLaunchKernelGGL(k_prep, dim3((unsigned)g), dim3(kBlk), 0, s, V, F, n, w.red);
    A3D_LAUNCH_CHECK();
  }
  if (n <= leaf_max) {
    hipLaunchKernelGGL(k_small_root, dim3(1), dim3(64), 0, s, V, F, n, nodes, tris, boxes);
    A3D_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_header, dim3(1), dim3(64), 0, s, (BvhHeader*)blob, n, cap, leaf_max, w.red, w.kept,
                       w.remap, (int64_t)0, (const int32_t*)nullptr);
    A3D_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_obbs, dim3(1), dim3(kBlk), 0, s, blob);
    A3D_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tri_obbs, dim3(1), dim3(kBlk), 0, s, blob);
    A3D_LAUNCH_CHECK();
    return A3D_OK;
  }
  unsigned gn = (unsigned)div_up(n, kBlk), gm = (unsigned)div_up(m, kBlk);
  hipLaunchKernelGGL(k_morton, dim3(gn), dim3(kBlk), 0, s, V, F, n, w.red, w.keys_in, w.vals_in);
  A3D_LAUNCH_CHECK();
  size_t cb = w.cub_bytes;
  A3D_HIP_CHECK(rocprim::radix_sort_pairs(w.cub, cb, w.keys_in, w.keys_out, w.vals_in, w.vals_out, (int)n, 0, 63, s));
  hipLaunchKernelGGL(k_karras, dim3(gm), dim3(kBlk), 0, s, w.keys_out, n, w.kl, w.kr, w.first, w.size, w.par_int, w.par_leaf, w.flag);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_leaf_boxes, dim3(gn), dim3(kBlk), 0, s, V, F, w.vals_out, n, w.leaf_box);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_walk_radius, dim3(1), dim3(64), 0, s, w.red);
  A3D_LAUNCH_CHECK();
  // Fit every internal node of the Karras tree and choose the leaves.  The
  // tree is at most kMaxDepth deep, so kMaxDepth + 1 rounds always finish
  // (the empty ones return at once) and no host synchronisation is needed.
  if ((rc = pass_init(w, m, true, s))) return rc;
  if ((rc = launch_rounds<false>(w, m, 0, kMaxDepth + 1, leaf_max, s))) return rc;
  hipLaunchKernelGGL(k_kept, dim3(gm), dim3(kBlk), 0, s, m, leaf_max, w.size, w.leafy, w.par_int, w.kept);
  A3D_LAUNCH_CHECK();
  const int32_t* depth_src = w.height;  // root = node 0
  if (treelet_passes > 0) {
    A3D_HIP_CHECK(hipMemcpyAsync(w.kl0, w.kl, m * 4, hipMemcpyDeviceToDevice, s));
    A3D_HIP_CHECK(hipMemcpyAsync(w.kr0, w.kr, m * 4, hipMemcpyDeviceToDevice, s));
    A3D_HIP_CHECK(hipMemcpyAsync(w.node_box0, w.node_box, m * 24, hipMemcpyDeviceToDevice, s));
    int32_t h0 = 0;
    A3D_HIP_CHECK(hipMemcpyAsync(&h0, w.height, 4, hipMemcpyDeviceToHost, s));
    for (int pass = 0; pass < treelet_passes; ++pass) {
      if ((rc = pass_init(w, m, false, s))) return rc;
      // Restructuring can deepen the tree, so the rounds run until the
      // frontier is empty, checked on the host every `chunk` rounds.
      const int chunk = 16;
      int r = 0;
      while (true) {
        if (r + chunk > kRoundsMax) return fail(A3D_ERR_INTERNAL, "a3d_bvh_build: treelet pass did not converge");
        if ((rc = launch_rounds<true>(w, m, r, chunk, leaf_max, s))) return rc;
        r += chunk;
        int32_t left = 0;
        A3D_HIP_CHECK(hipMemcpyAsync(&left, w.counts + r, 4, hipMemcpyDeviceToHost, s));
        A3D_HIP_CHECK(hipStreamSynchronize(s));
        if (left == 0) break;
      }
    }
    int32_t h1 = 0;
    A3D_HIP_CHECK(hipMemcpyAsync(&h1, w.height, 4, hipMemcpyDeviceToHost, s));
    A3D_HIP_CHECK(hipStreamSynchronize(s));
    if (h1 > kMaxDepth) {  // deeper than the walks' stacks allow: keep the fitted Karras tree
      A3D_HIP_CHECK(hipMemcpyAsync(w.kl, w.kl0, m * 4, hipMemcpyDeviceToDevice, s));
      A3D_HIP_CHECK(hipMemcpyAsync(w.kr, w.kr0, m * 4, hipMemcpyDeviceToDevice, s));
      A3D_HIP_CHECK(hipMemcpyAsync(w.node_box, w.node_box0, m * 24, hipMemcpyDeviceToDevice, s));
      A3D_HIP_CHECK(hipMemcpyAsync(w.height, &h0, 4, hipMemcpyHostToDevice, s));
      A3D_HIP_CHECK(hipStreamSynchronize(s));
    }
  }
  cb = w.cub_bytes;
  A3D_HIP_CHECK(a3d::exclusive_sum(w.cub, cb, w.kept, w.remap, (int)m, s));
  hipLaunchKernelGGL(k_emit_nodes, dim3(gm), dim3(kBlk), 0, s, m, w.kept, w.remap, w.kl, w.kr, w.size, w.first, w.leaf_box, w.node_box, nodes);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_emit_tris, dim3(gn), dim3(kBlk), 0, s, V, F, w.vals_out, n, tris, boxes);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_header, dim3(1), dim3(64), 0, s, (BvhHeader*)blob, n, cap, leaf_max, w.red, w.kept, w.remap, m,
                     depth_src);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_obbs, dim3((unsigned)div_up(2 * cap, kBlk)), dim3(kBlk), 0, s, blob);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_tri_obbs, dim3((unsigned)div_up(n, kBlk)), dim3(kBlk), 0, s, blob);
  A3D_LAUNCH_CHECK();
  // The walks' stacks hold trees up to kMaxWalkDepth internal levels; the
  // Karras tree of distinct keys stays within kMaxDepth < kMaxWalkDepth,
  // checked here rather than assumed.
  int32_t depth = 0;
  A3D_HIP_CHECK(hipMemcpyAsync(&depth, (const char*)blob + offsetof(BvhHeader, depth), 4, hipMemcpyDeviceToHost, s));
  A3D_HIP_CHECK(hipStreamSynchronize(s));
  if (depth > kMaxWalkDepth)
    return fail(A3D_ERR_CAPACITY, "a3d_bvh_build: tree depth %d exceeds the walks' stacks (%d)", depth, kMaxWalkDepth);
  return A3D_OK;
}

extern "C" int a3d_bvh_build(const float* V, int64_t nV, const int32_t* F, int64_t nF, int leaf_max, void* blob,
                             int64_t blob_bytes, void* ws, int64_t ws_bytes, a3d_stream_t stream) {
  return a3d_bvh_build_ex(V, nV, F, nF, leaf_max, A3D_BVH_TREELET_PASSES, blob, blob_bytes, ws, ws_bytes, stream);
}

extern "C" int a3d_bvh_info(const void* blob, int64_t* out4, a3d_stream_t stream) {
  if (!blob || !out4) return fail(A3D_ERR_ARG, "a3d_bvh_info: null argument");
  BvhHeader h;
  A3D_HIP_CHECK(hipMemcpyAsync(&h, blob, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream));
  A3D_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  if (h.magic != kBvhMagic) return fail(A3D_ERR_FORMAT, "a3d_bvh_info: not a BVH blob");
  out4[0] = h.n_tris;
  out4[1] = h.n_nodes;
  out4[2] = h.depth;
  out4[3] = h.leaf_max;
  return A3D_OK;
}
