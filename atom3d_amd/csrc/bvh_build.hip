// Device BVH build for gfx950: Morton-ordered linear BVH (Karras 2012,
// "Maximizing Parallelism in the Construction of BVHs, Octrees, and k-d
// Trees"), bottom-up box fitting with arrival counters, then collapse of
// subtrees of <= leaf_max triangles into leaves and emission of 64-byte
// two-child nodes.
//
// Replaces the reference's single-threaded host SAH build
// (Atom3D atom3d/kernels/bvh_kernels.cu:616-925), which copies the mesh to
// the host and back.  Query results never depend on the tree shape: the
// traversals in bvh_query.hip return the same faces for any valid tree.
#include <hipcub/hipcub.hpp>

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
enum { R_CLO = 0, R_CHI = 3, R_VLO = 6, R_VHI = 9, R_CMAX = 12, R_PAD = 13, R_N = 16 };

struct Ws {
  uint32_t* red;
  uint64_t* keys_in;
  uint64_t* keys_out;
  uint32_t* vals_in;
  uint32_t* vals_out;
  float* leaf_box;   // [n][6] sorted order
  float* node_box;   // [n-1][6]
  int32_t* kl;       // [n-1] child refs: >=0 internal, <0 ~leafprim
  int32_t* kr;
  int32_t* first;    // [n-1]
  int32_t* size;     // [n-1]
  int32_t* par_int;  // [n-1]
  int32_t* par_leaf; // [n]
  int32_t* flag;     // [n-1]
  int32_t* kept;     // [n-1]
  int32_t* remap;    // [n-1]
  void* cub;
  size_t cub_bytes;
  int64_t total;
};

inline int64_t al(int64_t x) { return (x + 255) & ~int64_t(255); }

size_t cub_bytes_for(int64_t n) {
  size_t sort_b = 0, scan_b = 0;
  int nn = (int)(n > 1 ? n : 2);
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_b, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                     (uint32_t*)nullptr, (uint32_t*)nullptr, nn, 0, 63);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_b, (int32_t*)nullptr, (int32_t*)nullptr, nn);
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
  w.cub_bytes = cub_bytes_for(n);
  w.cub = take((int64_t)w.cub_bytes);
  w.total = off;
  return w;
}

int64_t blob_size(int64_t nF) {
  int64_t cap = nF > 1 ? nF - 1 : 1;
  return 256 + 64 * cap + 48 * nF + 32 * nF;
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
// reductions.  One atomic per wave per slot.
__global__ __launch_bounds__(kBlk) void k_prep(const float* __restrict__ V, const int32_t* __restrict__ F,
                                                int64_t n, uint32_t* red) {
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
  int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < R_N; ++k) {
    if (k > R_PAD) continue;
    bool is_min = (k >= R_CLO && k < R_CHI) || (k >= R_VLO && k < R_VHI);
    if (k == R_CMAX) continue;
    float v = is_min ? wave_min(r[k]) : wave_max(r[k]);
    if (lane == 0) {
      if (is_min) atomicMin(&red[k], f2ord(v));
      else atomicMax(&red[k], f2ord(v));
    }
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

// Leaf boxes, then climb: the second child to arrive at a node fits its box.
__global__ __launch_bounds__(kBlk) void k_fit(const float* __restrict__ V, const int32_t* __restrict__ F,
                                               const uint32_t* __restrict__ order, int64_t n,
                                               const int32_t* __restrict__ kl, const int32_t* __restrict__ kr,
                                               const int32_t* __restrict__ par_int,
                                               const int32_t* __restrict__ par_leaf, int32_t* flag,
                                               float* leaf_box, float* node_box) {
  int64_t j = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (j >= n) return;
  float b[6];
  leaf_box_of(V, F, order[j], b);
  for (int k = 0; k < 6; ++k) leaf_box[j * 6 + k] = b[k];
  if (n == 1) return;
  int32_t p = par_leaf[j];
  while (p >= 0) {
    __threadfence();
    int old = atomicAdd(&flag[p], 1);
    if (old == 0) return;
    __threadfence();
    int32_t c[2] = {kl[p], kr[p]};
    float r[6] = {3.0e38f, 3.0e38f, 3.0e38f, -3.0e38f, -3.0e38f, -3.0e38f};
    for (int s = 0; s < 2; ++s) {
      const float* src = c[s] < 0 ? leaf_box + (int64_t)(~c[s]) * 6 : node_box + (int64_t)c[s] * 6;
      // The sibling's box was written by another CU (plain stores, then a
      // device-scope fence before its counter increment): read it with
      // device-scope loads so no stale per-CU cache line is used.
      float q[6];
      for (int k = 0; k < 6; ++k) q[k] = __hip_atomic_load(src + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int k = 0; k < 3; ++k) { r[k] = fminf(r[k], q[k]); r[k + 3] = fmaxf(r[k + 3], q[k + 3]); }
    }
    for (int k = 0; k < 6; ++k) node_box[(int64_t)p * 6 + k] = r[k];
    p = par_int[p];
  }
}

__global__ __launch_bounds__(kBlk) void k_kept(const int32_t* __restrict__ size, int64_t m, int leaf_max,
                                                int32_t* kept) {
  int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i >= m) return;
  kept[i] = size[i] > leaf_max ? 1 : 0;
}

__device__ __forceinline__ void child_info(int32_t c, const int32_t* __restrict__ size,
                                           const int32_t* __restrict__ first, const int32_t* __restrict__ remap,
                                           const float* __restrict__ leaf_box, const float* __restrict__ node_box,
                                           int leaf_max, int32_t& ref, const float*& box) {
  if (c < 0) {
    ref = leaf_ref(~c, 1);
    box = leaf_box + (int64_t)(~c) * 6;
  } else if (size[c] <= leaf_max) {
    ref = leaf_ref(first[c], size[c]);
    box = node_box + (int64_t)c * 6;
  } else {
    ref = remap[c];
    box = node_box + (int64_t)c * 6;
  }
}

__global__ __launch_bounds__(kBlk) void k_emit_nodes(int64_t m, int leaf_max, const int32_t* __restrict__ kept,
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
  child_info(kl[i], size, first, remap, leaf_box, node_box, leaf_max, lr, lb);
  child_info(kr[i], size, first, remap, leaf_box, node_box, leaf_max, rr, rb);
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

__global__ void k_header(BvhHeader* h, int64_t n, int64_t cap, int leaf_max, const uint32_t* __restrict__ red,
                         const int32_t* __restrict__ kept, const int32_t* __restrict__ remap, int64_t m) {
  if (threadIdx.x != 0) return;
  BvhHeader H;
  for (int k = 0; k < 44; ++k) H.pad[k] = 0;
  H.magic = kBvhMagic;
  H.n_tris = (int32_t)n;
  H.n_nodes = (n > leaf_max && m > 0) ? remap[m - 1] + kept[m - 1] : 1;
  H.n_node_cap = (int32_t)cap;
  H.nodes_off = 256;
  H.tris_off = 256 + 64 * cap;
  H.boxes_off = H.tris_off + 48 * n;
  float cm = 0.0f;
  for (int k = 0; k < 3; ++k) {
    H.scene_lo[k] = n > 0 ? ord2f(red[R_VLO + k]) : 0.0f;
    H.scene_hi[k] = n > 0 ? ord2f(red[R_VHI + k]) : 0.0f;
    cm = fmaxf(cm, fmaxf(fabsf(H.scene_lo[k]), fabsf(H.scene_hi[k])));
  }
  H.coord_max = cm;
  H.tri_pad_max = n > 0 ? ord2f(red[R_PAD]) : 0.0f;
  H.leaf_max = leaf_max;
  H.depth = -1;
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

extern "C" int a3d_bvh_build(const float* V, int64_t nV, const int32_t* F, int64_t nF, int leaf_max, void* blob,
                             int64_t blob_bytes, void* ws, int64_t ws_bytes, a3d_stream_t stream) {
  hipStream_t s = (hipStream_t)stream;
  int64_t need_blob = 0, need_ws = 0;
  int rc = a3d_bvh_sizes(nF, leaf_max, &need_blob, &need_ws);
  if (rc) return rc;
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
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_prep, dim3((unsigned)g), dim3(kBlk), 0, s, V, F, n, w.red);
    A3D_LAUNCH_CHECK();
  }
  if (n <= leaf_max) {
    hipLaunchKernelGGL(k_small_root, dim3(1), dim3(64), 0, s, V, F, n, nodes, tris, boxes);
    A3D_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_header, dim3(1), dim3(64), 0, s, (BvhHeader*)blob, n, cap, leaf_max, w.red, w.kept, w.remap, (int64_t)0);
    A3D_LAUNCH_CHECK();
    return A3D_OK;
  }
  unsigned gn = (unsigned)div_up(n, kBlk), gm = (unsigned)div_up(m, kBlk);
  hipLaunchKernelGGL(k_morton, dim3(gn), dim3(kBlk), 0, s, V, F, n, w.red, w.keys_in, w.vals_in);
  A3D_LAUNCH_CHECK();
  size_t cb = w.cub_bytes;
  A3D_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(w.cub, cb, w.keys_in, w.keys_out, w.vals_in, w.vals_out, (int)n, 0, 63, s));
  hipLaunchKernelGGL(k_karras, dim3(gm), dim3(kBlk), 0, s, w.keys_out, n, w.kl, w.kr, w.first, w.size, w.par_int, w.par_leaf, w.flag);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fit, dim3(gn), dim3(kBlk), 0, s, V, F, w.vals_out, n, w.kl, w.kr, w.par_int, w.par_leaf, w.flag, w.leaf_box, w.node_box);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_kept, dim3(gm), dim3(kBlk), 0, s, w.size, m, leaf_max, w.kept);
  A3D_LAUNCH_CHECK();
  cb = w.cub_bytes;
  A3D_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(w.cub, cb, w.kept, w.remap, (int)m, s));
  hipLaunchKernelGGL(k_emit_nodes, dim3(gm), dim3(kBlk), 0, s, m, leaf_max, w.kept, w.remap, w.kl, w.kr, w.size, w.first, w.leaf_box, w.node_box, nodes);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_emit_tris, dim3(gn), dim3(kBlk), 0, s, V, F, w.vals_out, n, tris, boxes);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_header, dim3(1), dim3(64), 0, s, (BvhHeader*)blob, n, cap, leaf_max, w.red, w.kept, w.remap, m);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
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
