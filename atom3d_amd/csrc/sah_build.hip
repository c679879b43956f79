// The reference's binned-SAH tree, rebuilt on the device, for the box
// queries (SAT intersect, octree level test, voxelizer).
//
// The reference keeps a (box, face) pair when the box overlaps the face's
// LEAF box in its host-built SAH tree and the face passes the SAT test
// (Atom3D atom3d/kernels/bvh_kernels.cu:583-591; tree from
// build_bvh_recursive :667-854, 4 triangles per leaf by default,
// kernels/bvh.py:66-93).  The SAT test's 1e-6 slack accepts faces whose own
// box misses the query box, so the kept set depends on which faces share a
// leaf: this file builds exactly the reference's tree (same leaf sets, same
// node boxes, same triangle order) and the box walks run on it.
//
// Why the result is the reference's, step by step (per node [start, end)):
//   * node box, centroid bounds, SAH bins: minima, maxima and counts over
//     the node's triangles — independent of the order they are combined in,
//     so the level passes reduce them with atomics;
//   * split choice: one thread evaluates the 3 x 15 candidate costs from the
//     bins in the reference's order and operations (compiled with
//     -ffp-contract=off, IEEE division), first strict minimum;
//   * partition on centroid < split: the reference calls std::partition,
//     which in libstdc++ (GCC 11 bits/stl_algo.h, bidirectional __partition)
//     swaps the k-th element failing the predicate from the left with the
//     k-th element passing it from the right.  The level passes compute both
//     ranks with prefix sums and swap the pairs, giving the same order;
//   * median fallbacks (empty SAH partition, or no axis to bin): the
//     reference calls std::nth_element; libstdc++'s introselect
//     (median-of-3 pivot, unguarded Hoare partition, heap select after
//     2 floor(log2 n) rounds, insertion sort below 4 elements) is restated in
//     seq_nth_element and run by one thread on the node's range.  Ties make
//     nth_element order dependent; since every earlier step reproduces the
//     reference's order, so does this one;
//   * nodes of at most kSmall triangles are built by one thread each with a
//     sequential restatement of the recursion (the same partition and
//     fallback code), depth first.
// Node numbering differs from the reference's (level order, then one block
// of records per small subtree); no query result depends on it.
//
// Blob: the BVH blob layout of a3d_common.hpp (header, Node64 records,
// Tri48 triangles in the reference's triangle order), without the UDF
// walk's padded and oriented boxes (boxes_off = obbs_off = tobbs_off = 0);
// header.pad[0] = kBlobSah.
#include "a3d_prim.hpp"

#include <cstddef>

#include "a3d_common.hpp"

namespace a3d {
namespace {

constexpr int kBlk = 256;
constexpr int kChunk = 4096;  // positions per block in the level passes (16 per thread)
constexpr int kRounds = kChunk / kBlk;
constexpr int kSmall = 64;    // nodes of at most this many triangles: one thread builds the subtree
constexpr int kBins = 16;     // bvh_kernels.cu:704
constexpr int kBinWords = 7;  // count, lo[3], hi[3]
constexpr int kNodeBins = 3 * kBins * kBinWords;
constexpr int32_t kBlobSah = 1;

// Per triangle (original order): vertex box and centroid (lo.w = cx, hi.w =
// cy; cz separately), computed as build_bvh_cuda does (bvh_kernels.cu:875-889).
struct SahTri {
  float4 lo;
  float4 hi;
};

__device__ __forceinline__ uint32_t f2ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ float ckey(const SahTri* __restrict__ tri, const float* __restrict__ cz, int32_t t,
                                      int ax) {
  return ax == 0 ? tri[t].lo.w : (ax == 1 ? tri[t].hi.w : cz[t]);
}

// Level node: triangle range, own record, parent slot (record * 2 + side,
// -1 for the root), chunk range and depth.
struct LNode {
  int32_t start, end, rec, prec, c0, nc, depth, pad;
};
// Subtree of at most kSmall triangles, built by one thread.
struct SNode {
  int32_t start, end, prec, depth;
};

enum { C_NB = 0, C_NC = 1, C_NS = 2, C_NTH = 3, C_NREC = 4, C_DEPTH = 5, C_ERR = 6, C_RED = 8, C_N = 32 };

// Writes a child's box and reference into its parent's record.
__device__ __forceinline__ void put_child(Node64* __restrict__ nodes, int32_t prec, const float* lo, const float* hi,
                                          int32_t ref) {
  if (prec < 0) return;
  float* r = reinterpret_cast<float*>(nodes + (prec >> 1));
  const int side = prec & 1;
  r[side * 6 + 0] = lo[0];
  r[side * 6 + 1] = lo[1];
  r[side * 6 + 2] = lo[2];
  r[side * 6 + 3] = hi[0];
  r[side * 6 + 4] = hi[1];
  r[side * 6 + 5] = hi[2];
  int32_t* d = reinterpret_cast<int32_t*>(r + 12);
  d[side] = ref;
  if (side == 0) {
    d[2] = 0;
    d[3] = 0;
  }
}

// ------------------------------------------------------------------ shared
// SAHBin::area (bvh_kernels.cu:652-662).
__device__ __forceinline__ float bin_area(int count, const float* lo, const float* hi) {
  if (count == 0) return 0.0f;
  float dx = hi[0] - lo[0];
  float dy = hi[1] - lo[1];
  float dz = hi[2] - lo[2];
  dx = fmaxf(0.0f, dx);
  dy = fmaxf(0.0f, dy);
  dz = fmaxf(0.0f, dz);
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// The cost sweep of one axis (bvh_kernels.cu:744-776) over 16 bins given as
// counts and boxes (empty bins: lo 1e30, hi -1e30, as SAHBin's constructor).
__device__ __forceinline__ void axis_sweep(const int* cnt, const float* blo, const float* bhi, int axis,
                                           float& best_cost, int& best_axis, int& best_split) {
  float la[kBins - 1], ra[kBins - 1];
  int lc[kBins - 1], rc[kBins - 1];
  float clo[3] = {1e30f, 1e30f, 1e30f}, chi[3] = {-1e30f, -1e30f, -1e30f};
  int cc = 0;
  for (int i = 0; i < kBins - 1; i++) {
    for (int k = 0; k < 3; ++k) {
      clo[k] = fminf(clo[k], blo[i * 3 + k]);
      chi[k] = fmaxf(chi[k], bhi[i * 3 + k]);
    }
    cc += cnt[i];
    la[i] = bin_area(cc, clo, chi);
    lc[i] = cc;
  }
  for (int k = 0; k < 3; ++k) {
    clo[k] = 1e30f;
    chi[k] = -1e30f;
  }
  cc = 0;
  for (int i = kBins - 1; i > 0; i--) {
    for (int k = 0; k < 3; ++k) {
      clo[k] = fminf(clo[k], blo[i * 3 + k]);
      chi[k] = fmaxf(chi[k], bhi[i * 3 + k]);
    }
    cc += cnt[i];
    ra[i - 1] = bin_area(cc, clo, chi);
    rc[i - 1] = cc;
  }
  for (int i = 0; i < kBins - 1; i++) {
    float cost = lc[i] * la[i] + rc[i] * ra[i];
    if (cost < best_cost) {
      best_cost = cost;
      best_axis = axis;
      best_split = i;
    }
  }
}

// bvh_kernels.cu:790 (split position) and :723-737 (scale, bin index).
__device__ __forceinline__ float split_value(float lo, float hi, int best_split) {
  return lo + (best_split + 1) * (hi - lo) / kBins;
}
__device__ __forceinline__ float bin_scale(float lo, float hi) {
  float scale = kBins / (hi - lo);
  scale *= 0.999f;
  return scale;
}
__device__ __forceinline__ int bin_of(float c, float lo, float scale) {
  int b = (int)((c - lo) * scale);
  return b < 0 ? 0 : (b > kBins - 1 ? kBins - 1 : b);
}
// Longest axis, first on ties (bvh_kernels.cu:823-826).
__device__ __forceinline__ int fallback_axis(const float* dims) {
  int fa = 0;
  if (dims[1] > dims[0]) fa = 1;
  if (dims[2] > dims[fa]) fa = 2;
  return fa;
}

// ------------------------------------------------------------------ sequential restatements
// libstdc++ std::partition for random-access iterators (bidirectional
// __partition): returns the first position failing the predicate.
__device__ int seq_partition(int32_t* __restrict__ perm, int32_t first, int32_t last, const SahTri* __restrict__ tri,
                             const float* __restrict__ cz, int ax, float split) {
  while (true) {
    while (true) {
      if (first == last) return first;
      if (ckey(tri, cz, perm[first], ax) < split) ++first;
      else break;
    }
    --last;
    while (true) {
      if (first == last) return first;
      if (!(ckey(tri, cz, perm[last], ax) < split)) --last;
      else break;
    }
    const int32_t t = perm[first];
    perm[first] = perm[last];
    perm[last] = t;
    ++first;
  }
}

// std::nth_element(first, nth, last, comp) with comp(a, b) = key(a) < key(b),
// restating libstdc++ (GCC 11 bits/stl_algo.h __introselect and helpers,
// bits/stl_heap.h __make_heap / __adjust_heap / __push_heap / __pop_heap).
struct NthCtx {
  int32_t* perm;
  const SahTri* tri;
  const float* cz;
  int ax;
  __device__ float key(int32_t pos) const { return ckey(tri, cz, perm[pos], ax); }
  __device__ float keyv(int32_t id) const { return ckey(tri, cz, id, ax); }
  __device__ bool lt(int32_t i, int32_t j) const { return key(i) < key(j); }
  __device__ void swap(int32_t i, int32_t j) const {
    const int32_t t = perm[i];
    perm[i] = perm[j];
    perm[j] = t;
  }
};

__device__ void push_heap_(const NthCtx& c, int32_t first, int32_t hole, int32_t top, int32_t value) {
  int32_t parent = (hole - 1) / 2;
  while (hole > top && c.key(first + parent) < c.keyv(value)) {
    c.perm[first + hole] = c.perm[first + parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  c.perm[first + hole] = value;
}

__device__ void adjust_heap_(const NthCtx& c, int32_t first, int32_t hole, int32_t len, int32_t value) {
  const int32_t top = hole;
  int32_t second = hole;
  while (second < (len - 1) / 2) {
    second = 2 * (second + 1);
    if (c.lt(first + second, first + (second - 1))) second--;
    c.perm[first + hole] = c.perm[first + second];
    hole = second;
  }
  if ((len & 1) == 0 && second == (len - 2) / 2) {
    second = 2 * (second + 1);
    c.perm[first + hole] = c.perm[first + (second - 1)];
    hole = second - 1;
  }
  push_heap_(c, first, hole, top, value);
}

__device__ void heap_select_(const NthCtx& c, int32_t first, int32_t middle, int32_t last) {
  const int32_t len = middle - first;
  if (len >= 2) {  // __make_heap
    int32_t parent = (len - 2) / 2;
    while (true) {
      adjust_heap_(c, first, parent, len, c.perm[first + parent]);
      if (parent == 0) break;
      parent--;
    }
  }
  for (int32_t i = middle; i < last; ++i) {
    if (c.lt(i, first)) {  // __pop_heap(first, middle, i)
      const int32_t value = c.perm[i];
      c.perm[i] = c.perm[first];
      adjust_heap_(c, first, 0, len, value);
    }
  }
}

__device__ void insertion_sort_(const NthCtx& c, int32_t first, int32_t last) {
  if (first == last) return;
  for (int32_t i = first + 1; i != last; ++i) {
    if (c.lt(i, first)) {
      const int32_t val = c.perm[i];
      for (int32_t k = i; k > first; --k) c.perm[k] = c.perm[k - 1];
      c.perm[first] = val;
    } else {  // __unguarded_linear_insert
      const int32_t val = c.perm[i];
      const float kv = c.keyv(val);
      int32_t hole = i, next = i - 1;
      while (kv < c.key(next)) {
        c.perm[hole] = c.perm[next];
        hole = next;
        --next;
      }
      c.perm[hole] = val;
    }
  }
}

__device__ void move_median_to_first_(const NthCtx& c, int32_t result, int32_t a, int32_t b, int32_t cc) {
  if (c.lt(a, b)) {
    if (c.lt(b, cc)) c.swap(result, b);
    else if (c.lt(a, cc)) c.swap(result, cc);
    else c.swap(result, a);
  } else if (c.lt(a, cc)) c.swap(result, a);
  else if (c.lt(b, cc)) c.swap(result, cc);
  else c.swap(result, b);
}

__device__ int32_t unguarded_partition_pivot_(const NthCtx& c, int32_t first, int32_t last) {
  const int32_t mid = first + (last - first) / 2;
  move_median_to_first_(c, first, first + 1, mid, last - 1);
  const int32_t pivot = first;
  ++first;
  while (true) {  // __unguarded_partition(first + 1, last, pivot)
    while (c.lt(first, pivot)) ++first;
    --last;
    while (c.lt(pivot, last)) --last;
    if (!(first < last)) return first;
    c.swap(first, last);
    ++first;
  }
}

__device__ void seq_nth_element(int32_t* perm, const SahTri* tri, const float* cz, int ax, int32_t first,
                                int32_t nth, int32_t last) {
  if (first == last || nth == last) return;
  NthCtx c{perm, tri, cz, ax};
  int32_t depth_limit = 2 * (31 - __builtin_clz((uint32_t)(last - first)));  // 2 * std::__lg(n)
  while (last - first > 3) {
    if (depth_limit == 0) {
      heap_select_(c, first, nth + 1, last);
      c.swap(first, nth);
      return;
    }
    --depth_limit;
    const int32_t cut = unguarded_partition_pivot_(c, first, last);
    if (cut <= nth) first = cut;
    else last = cut;
  }
  insertion_sort_(c, first, last);
}

__device__ __forceinline__ void seq_box(const int32_t* __restrict__ perm, const SahTri* __restrict__ tri, int32_t s,
                                        int32_t e, float* lo, float* hi) {
  lo[0] = lo[1] = lo[2] = 1e30f;
  hi[0] = hi[1] = hi[2] = -1e30f;
  for (int32_t i = s; i < e; ++i) {
    const SahTri t = tri[perm[i]];
    lo[0] = fminf(lo[0], t.lo.x);
    lo[1] = fminf(lo[1], t.lo.y);
    lo[2] = fminf(lo[2], t.lo.z);
    hi[0] = fmaxf(hi[0], t.hi.x);
    hi[1] = fmaxf(hi[1], t.hi.y);
    hi[2] = fmaxf(hi[2], t.hi.z);
  }
}

// ------------------------------------------------------------------ kernels
__global__ void k_sah_init(int32_t* __restrict__ ctr) {
  const int i = threadIdx.x;
  if (i < C_N) ctr[i] = (i >= C_RED && i < C_RED + 3) ? (int32_t)0xffffffffu : 0;  // ordered min identity
}

// Per-triangle box and centroid; identity permutation; scene box.
__global__ __launch_bounds__(kBlk) void k_sah_prep(const float* __restrict__ V, const int32_t* __restrict__ F,
                                                    int64_t n, SahTri* __restrict__ tri, float* __restrict__ cz,
                                                    int32_t* __restrict__ perm, int32_t* __restrict__ ctr) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlk) {
    const int64_t i0 = F[t * 3], i1 = F[t * 3 + 1], i2 = F[t * 3 + 2];
    const v3 a = mk(V[i0 * 3], V[i0 * 3 + 1], V[i0 * 3 + 2]);
    const v3 b = mk(V[i1 * 3], V[i1 * 3 + 1], V[i1 * 3 + 2]);
    const v3 c = mk(V[i2 * 3], V[i2 * 3 + 1], V[i2 * 3 + 2]);
    SahTri s;
    s.lo = make_float4(fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z),
                       (a.x + b.x + c.x) / 3.0f);
    s.hi = make_float4(fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z),
                       (a.y + b.y + c.y) / 3.0f);
    tri[t] = s;
    cz[t] = (a.z + b.z + c.z) / 3.0f;
    perm[t] = (int32_t)t;
    lo[0] = fminf(lo[0], s.lo.x);
    lo[1] = fminf(lo[1], s.lo.y);
    lo[2] = fminf(lo[2], s.lo.z);
    hi[0] = fmaxf(hi[0], s.hi.x);
    hi[1] = fmaxf(hi[1], s.hi.y);
    hi[2] = fmaxf(hi[2], s.hi.z);
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 3; ++k) {
      lo[k] = fminf(lo[k], __shfl_xor(lo[k], off, 64));
      hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off, 64));
    }
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 3; ++k) {
      atomicMin((uint32_t*)&ctr[C_RED + k], f2ord(lo[k]));
      atomicMax((uint32_t*)&ctr[C_RED + 3 + k], f2ord(hi[k]));
    }
}

// Resets the reductions of nb level nodes.
__global__ __launch_bounds__(kBlk) void k_sah_level_init(int64_t nb, uint32_t* __restrict__ acc,
                                                          uint32_t* __restrict__ bins) {
  const int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i < nb * 12) acc[i] = (i % 12) % 6 < 3 ? 0xffffffffu : 0u;  // box lo, hi, centroid lo, hi
  if (i < nb * kNodeBins) {
    const int w = (int)(i % kBinWords);
    bins[i] = w == 0 ? 0u : (w < 4 ? 0xffffffffu : 0u);
  }
}

__device__ __forceinline__ void chunk_range(const int2* __restrict__ chunks, const LNode* __restrict__ nodes,
                                            int64_t c, int32_t& node, int32_t& lo, int32_t& hi) {
  const int2 ch = chunks[c];
  node = ch.x;
  lo = ch.y;
  const int32_t e = nodes[node].end;
  hi = lo + kChunk < e ? lo + kChunk : e;
}

// Node box and centroid bounds (bvh_kernels.cu:675-684, 711-720).
__global__ __launch_bounds__(kBlk) void k_sah_stats(const int2* __restrict__ chunks, const LNode* __restrict__ nodes,
                                                     const int32_t* __restrict__ perm, const SahTri* __restrict__ tri,
                                                     const float* __restrict__ cz, uint32_t* __restrict__ acc) {
  int32_t node, lo, hi;
  chunk_range(chunks, nodes, blockIdx.x, node, lo, hi);
  float v[12];
  for (int k = 0; k < 6; ++k) {
    v[k] = k % 6 < 3 ? INFINITY : -INFINITY;
    v[6 + k] = k % 6 < 3 ? INFINITY : -INFINITY;
  }
  for (int r = 0; r < kRounds; ++r) {
    const int32_t p = lo + r * kBlk + threadIdx.x;
    if (p >= hi) break;
    const int32_t t = perm[p];
    const SahTri s = tri[t];
    const float z = cz[t];
    v[0] = fminf(v[0], s.lo.x);
    v[1] = fminf(v[1], s.lo.y);
    v[2] = fminf(v[2], s.lo.z);
    v[3] = fmaxf(v[3], s.hi.x);
    v[4] = fmaxf(v[4], s.hi.y);
    v[5] = fmaxf(v[5], s.hi.z);
    v[6] = fminf(v[6], s.lo.w);
    v[7] = fminf(v[7], s.hi.w);
    v[8] = fminf(v[8], z);
    v[9] = fmaxf(v[9], s.lo.w);
    v[10] = fmaxf(v[10], s.hi.w);
    v[11] = fmaxf(v[11], z);
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 12; ++k) {
      const float o = __shfl_xor(v[k], off, 64);
      v[k] = (k % 6) < 3 ? fminf(v[k], o) : fmaxf(v[k], o);
    }
  __shared__ float s_v[kBlk / 64][12];
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 12; ++k) s_v[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 12) {
    const int k = threadIdx.x;
    float x = s_v[0][k];
    for (int w = 1; w < kBlk / 64; ++w) x = (k % 6) < 3 ? fminf(x, s_v[w][k]) : fmaxf(x, s_v[w][k]);
    if ((k % 6) < 3) atomicMin(&acc[node * 12 + k], f2ord(x));
    else atomicMax(&acc[node * 12 + k], f2ord(x));
  }
}

struct NodeStats {
  float blo[3], bhi[3], clo[3], chi[3];
};
__device__ __forceinline__ NodeStats load_stats(const uint32_t* __restrict__ acc, int32_t node) {
  NodeStats s;
  for (int k = 0; k < 3; ++k) {
    s.blo[k] = ord2f(acc[node * 12 + k]);
    s.bhi[k] = ord2f(acc[node * 12 + 3 + k]);
    s.clo[k] = ord2f(acc[node * 12 + 6 + k]);
    s.chi[k] = ord2f(acc[node * 12 + 9 + k]);
  }
  return s;
}
// Axes the reference bins: box extent and centroid extent both >= 1e-6
// (bvh_kernels.cu:706-721).
__device__ __forceinline__ bool axis_binned(const NodeStats& s, int ax) {
  if (s.bhi[ax] - s.blo[ax] < 1e-6f) return false;
  if (s.chi[ax] - s.clo[ax] < 1e-6f) return false;
  return true;
}

// SAH bins (bvh_kernels.cu:723-737): block-private LDS bins, then one global
// atomic per non-empty bin word.
__global__ __launch_bounds__(kBlk) void k_sah_bins(const int2* __restrict__ chunks, const LNode* __restrict__ nodes,
                                                    const int32_t* __restrict__ perm, const SahTri* __restrict__ tri,
                                                    const float* __restrict__ cz, const uint32_t* __restrict__ acc,
                                                    int leaf_n, uint32_t* __restrict__ bins) {
  __shared__ uint32_t s_bin[3 * kBins * kBinWords];
  int32_t node, lo, hi;
  chunk_range(chunks, nodes, blockIdx.x, node, lo, hi);
  if (nodes[node].end - nodes[node].start <= leaf_n) return;
  for (int i = threadIdx.x; i < 3 * kBins * kBinWords; i += kBlk) {
    const int w = i % kBinWords;
    s_bin[i] = w == 0 ? 0u : (w < 4 ? 0xffffffffu : 0u);
  }
  const NodeStats st = load_stats(acc, node);
  bool ok[3];
  float scale[3];
  for (int ax = 0; ax < 3; ++ax) {
    ok[ax] = axis_binned(st, ax);
    scale[ax] = ok[ax] ? bin_scale(st.clo[ax], st.chi[ax]) : 0.0f;
  }
  __syncthreads();
  for (int r = 0; r < kRounds; ++r) {
    const int32_t p = lo + r * kBlk + threadIdx.x;
    if (p >= hi) break;
    const int32_t t = perm[p];
    const SahTri s = tri[t];
    const float c[3] = {s.lo.w, s.hi.w, cz[t]};
    for (int ax = 0; ax < 3; ++ax) {
      if (!ok[ax]) continue;
      uint32_t* b = s_bin + (ax * kBins + bin_of(c[ax], st.clo[ax], scale[ax])) * kBinWords;
      atomicAdd(&b[0], 1u);
      atomicMin(&b[1], f2ord(s.lo.x));
      atomicMin(&b[2], f2ord(s.lo.y));
      atomicMin(&b[3], f2ord(s.lo.z));
      atomicMax(&b[4], f2ord(s.hi.x));
      atomicMax(&b[5], f2ord(s.hi.y));
      atomicMax(&b[6], f2ord(s.hi.z));
    }
  }
  __syncthreads();
  uint32_t* g = bins + (int64_t)node * kNodeBins;
  for (int i = threadIdx.x; i < 3 * kBins * kBinWords; i += kBlk) {
    const int w = i % kBinWords;
    if (s_bin[i - w] == 0) continue;
    if (w == 0) atomicAdd(&g[i], s_bin[i]);
    else if (w < 4) atomicMin(&g[i], s_bin[i]);
    else atomicMax(&g[i], s_bin[i]);
  }
}

// Split choice per level node (bvh_kernels.cu:686-838); also writes the
// node's box and record into its parent's slot.  mode 0: partition on
// centroid[axis] < split; mode 1: median on axis (no axis binned).
__global__ __launch_bounds__(kBlk) void k_sah_split(int64_t nb, const LNode* __restrict__ nodes,
                                                     const uint32_t* __restrict__ acc,
                                                     const uint32_t* __restrict__ bins, Node64* __restrict__ out,
                                                     int32_t* __restrict__ axis_o, float* __restrict__ split_o,
                                                     int32_t* __restrict__ mode_o) {
  const int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i >= nb) return;
  const LNode nd = nodes[i];
  const NodeStats st = load_stats(acc, (int32_t)i);
  put_child(out, nd.prec, st.blo, st.bhi, nd.rec);
  float best_cost = 1e30f;
  int best_axis = -1, best_split = -1;
  const uint32_t* g = bins + i * kNodeBins;
  for (int ax = 0; ax < 3; ++ax) {
    if (!axis_binned(st, ax)) continue;
    int cnt[kBins];
    float blo[kBins * 3], bhi[kBins * 3];
    for (int b = 0; b < kBins; ++b) {
      const uint32_t* w = g + (ax * kBins + b) * kBinWords;
      cnt[b] = (int)w[0];
      for (int k = 0; k < 3; ++k) {
        blo[b * 3 + k] = cnt[b] ? ord2f(w[1 + k]) : 1e30f;
        bhi[b * 3 + k] = cnt[b] ? ord2f(w[4 + k]) : -1e30f;
      }
    }
    axis_sweep(cnt, blo, bhi, ax, best_cost, best_axis, best_split);
  }
  if (best_axis != -1) {
    axis_o[i] = best_axis;
    split_o[i] = split_value(st.clo[best_axis], st.chi[best_axis], best_split);
    mode_o[i] = 0;
  } else {
    const float dims[3] = {st.bhi[0] - st.blo[0], st.bhi[1] - st.blo[1], st.bhi[2] - st.blo[2]};
    axis_o[i] = fallback_axis(dims);
    split_o[i] = 0.0f;
    mode_o[i] = 1;
  }
}

// Block-wide exclusive prefix of a predicate over one round of kBlk
// positions; returns this thread's prefix and adds the round total to run.
__device__ __forceinline__ int32_t round_prefix(bool pred, int32_t* s_w, int32_t& run) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t m = __ballot(pred);
  const int32_t in_wave = __popcll(m & ((lane ? (~0ull >> (64 - lane)) : 0ull)));
  if (lane == 0) s_w[wave] = __popcll(m);
  __syncthreads();
  int32_t before = run;
  for (int w = 0; w < wave; ++w) before += s_w[w];
  int32_t tot = 0;
  for (int w = 0; w < kBlk / 64; ++w) tot += s_w[w];
  __syncthreads();
  run += tot;
  return before + in_wave;
}

__global__ __launch_bounds__(kBlk) void k_sah_count(int64_t nc, const int2* __restrict__ chunks,
                                                     const LNode* __restrict__ nodes, const int32_t* __restrict__ perm,
                                                     const SahTri* __restrict__ tri, const float* __restrict__ cz,
                                                     const int32_t* __restrict__ axis_a,
                                                     const float* __restrict__ split_a,
                                                     const int32_t* __restrict__ mode_a, int leaf_n,
                                                     int32_t* __restrict__ chunk_true) {
  int32_t node, lo, hi;
  chunk_range(chunks, nodes, blockIdx.x, node, lo, hi);
  if (blockIdx.x == 0 && threadIdx.x == 0) chunk_true[nc] = 0;
  __shared__ int32_t s_c[kBlk / 64];
  if (threadIdx.x < kBlk / 64) s_c[threadIdx.x] = 0;
  __syncthreads();
  const bool live = nodes[node].end - nodes[node].start > leaf_n && mode_a[node] == 0;
  int32_t c = 0;
  if (live) {
    const int ax = axis_a[node];
    const float sp = split_a[node];
    for (int r = 0; r < kRounds; ++r) {
      const int32_t p = lo + r * kBlk + threadIdx.x;
      if (p >= hi) break;
      c += ckey(tri, cz, perm[p], ax) < sp;
    }
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63) == 0) s_c[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t t = 0;
    for (int w = 0; w < kBlk / 64; ++w) t += s_c[w];
    chunk_true[blockIdx.x] = t;
  }
}

// Pair ranks of the two-pointer partition: the j-th position failing the
// predicate in [start, start + L) and the j-th passing one counted from the
// right in [start + L, end) swap (S pairs).
__global__ __launch_bounds__(kBlk) void k_sah_rank(const int2* __restrict__ chunks, const LNode* __restrict__ nodes,
                                                    const int32_t* __restrict__ perm, const SahTri* __restrict__ tri,
                                                    const float* __restrict__ cz, const int32_t* __restrict__ axis_a,
                                                    const float* __restrict__ split_a,
                                                    const int32_t* __restrict__ mode_a,
                                                    const int32_t* __restrict__ chunk_pre, int leaf_n,
                                                    int32_t* __restrict__ pos_lf, int32_t* __restrict__ pos_rt,
                                                    int32_t* __restrict__ swaps) {
  __shared__ int32_t s_w[kBlk / 64];
  int32_t node, lo, hi;
  chunk_range(chunks, nodes, blockIdx.x, node, lo, hi);
  const LNode nd = nodes[node];
  const int32_t cnt = nd.end - nd.start;
  if (cnt <= leaf_n || mode_a[node] != 0) return;
  const int32_t L = chunk_pre[nd.c0 + nd.nc] - chunk_pre[nd.c0];
  if (L == 0 || L == cnt) return;  // empty partition: median fallback
  const int ax = axis_a[node];
  const float sp = split_a[node];
  int32_t run = chunk_pre[blockIdx.x] - chunk_pre[nd.c0];
  const int32_t mid = nd.start + L;
  for (int r = 0; r < kRounds; ++r) {
    if (lo + r * kBlk >= hi) break;  // uniform over the block
    const int32_t p = lo + r * kBlk + threadIdx.x;
    const bool in = p < hi;
    const bool pred = in && ckey(tri, cz, perm[p], ax) < sp;
    const int32_t T = round_prefix(pred, s_w, run);  // passing positions in [start, p)
    if (!in) continue;
    if (p < mid && !pred) pos_lf[nd.start + (p - nd.start) - T] = p;
    if (p >= mid && pred) pos_rt[nd.start + (L - T - 1)] = p;
    if (p == mid - 1) swaps[node] = L - (T + (pred ? 1 : 0));
  }
}

__global__ __launch_bounds__(kBlk) void k_sah_swap(const int2* __restrict__ chunks, const LNode* __restrict__ nodes,
                                                    const int32_t* __restrict__ mode_a,
                                                    const int32_t* __restrict__ chunk_pre, int leaf_n,
                                                    const int32_t* __restrict__ pos_lf,
                                                    const int32_t* __restrict__ pos_rt,
                                                    const int32_t* __restrict__ swaps, int32_t* __restrict__ perm) {
  int32_t node, lo, hi;
  chunk_range(chunks, nodes, blockIdx.x, node, lo, hi);
  const LNode nd = nodes[node];
  const int32_t cnt = nd.end - nd.start;
  if (cnt <= leaf_n || mode_a[node] != 0) return;
  const int32_t L = chunk_pre[nd.c0 + nd.nc] - chunk_pre[nd.c0];
  if (L == 0 || L == cnt) return;
  const int32_t S = swaps[node];
  for (int r = 0; r < kRounds; ++r) {
    const int32_t p = lo + r * kBlk + threadIdx.x;
    if (p >= hi) break;
    const int32_t j = p - nd.start;
    if (j >= S) break;
    const int32_t a = pos_lf[nd.start + j], b = pos_rt[nd.start + j];
    const int32_t ta = perm[a];
    perm[a] = perm[b];
    perm[b] = ta;
  }
}

enum { K_LEAF = 0, K_SMALL = 1, K_BIG = 2 };

// Children of each level node; median nodes are queued for k_sah_nth.
__global__ __launch_bounds__(kBlk) void k_sah_children(int64_t nb, const LNode* __restrict__ nodes,
                                                        const int32_t* __restrict__ axis_a,
                                                        const int32_t* __restrict__ mode_a,
                                                        const int32_t* __restrict__ chunk_pre, int leaf_n,
                                                        int4* __restrict__ child, int32_t* __restrict__ f_big,
                                                        int32_t* __restrict__ f_small, int32_t* __restrict__ f_chunk,
                                                        int32_t* __restrict__ nth_list, int32_t* __restrict__ ctr) {
  const int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i >= nb) return;
  const LNode nd = nodes[i];
  const int32_t cnt = nd.end - nd.start;
  int32_t mid = nd.start + cnt / 2;
  bool median = true;
  if (mode_a[i] == 0) {
    const int32_t L = chunk_pre[nd.c0 + nd.nc] - chunk_pre[nd.c0];
    if (L != 0 && L != cnt) {
      mid = nd.start + L;
      median = false;
    }
  }
  if (median) nth_list[atomicAdd(&ctr[C_NTH], 1)] = (int32_t)i;
  for (int side = 0; side < 2; ++side) {
    const int32_t s = side ? mid : nd.start, e = side ? nd.end : mid;
    const int32_t c = e - s;
    const int kind = c <= leaf_n ? K_LEAF : (c <= kSmall ? K_SMALL : K_BIG);
    child[2 * i + side] = make_int4(s, e, nd.rec * 2 + side, nd.depth + 1);
    f_big[2 * i + side] = kind == K_BIG;
    f_small[2 * i + side] = kind == K_SMALL;
    f_chunk[2 * i + side] = kind == K_BIG ? (c + kChunk - 1) / kChunk : 0;
  }
  (void)axis_a;
}

__global__ void k_sah_nth(const LNode* __restrict__ nodes, const int32_t* __restrict__ nth_list,
                          const int32_t* __restrict__ axis_a, const int32_t* __restrict__ ctr,
                          int32_t* __restrict__ perm, const SahTri* __restrict__ tri, const float* __restrict__ cz) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ctr[C_NTH]) return;
  const int32_t i = nth_list[k];
  const LNode nd = nodes[i];
  seq_nth_element(perm, tri, cz, axis_a[i], nd.start, nd.start + (nd.end - nd.start) / 2, nd.end);
}

// Next level's nodes and chunk table, the small-subtree list, and the leaf
// children's records.
__global__ __launch_bounds__(kBlk) void k_sah_emit(int64_t nch, const int4* __restrict__ child,
                                                    const int32_t* __restrict__ f_big,
                                                    const int32_t* __restrict__ f_small,
                                                    const int32_t* __restrict__ f_chunk,
                                                    const int32_t* __restrict__ o_big,
                                                    const int32_t* __restrict__ o_small,
                                                    const int32_t* __restrict__ o_chunk, int32_t rec_next,
                                                    int32_t small_base, int64_t nb_max, int64_t nc_max,
                                                    LNode* __restrict__ next,
                                                    int2* __restrict__ chunks_next, SNode* __restrict__ small,
                                                    const int32_t* __restrict__ perm, const SahTri* __restrict__ tri,
                                                    Node64* __restrict__ out, int32_t* __restrict__ ctr) {
  const int64_t c = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (c >= nch) return;
  const int4 ch = child[c];
  if (f_big[c]) {
    const int32_t j = o_big[c];
    if (j >= nb_max || o_chunk[c] + f_chunk[c] > nc_max) {  // cannot happen (every big node holds > kSmall)
      atomicOr(&ctr[C_ERR], 1);
      return;
    }
    LNode nd;
    nd.start = ch.x;
    nd.end = ch.y;
    nd.rec = rec_next + j;
    nd.prec = ch.z;
    nd.c0 = o_chunk[c];
    nd.nc = f_chunk[c];
    nd.depth = ch.w;
    nd.pad = 0;
    next[j] = nd;
    for (int32_t k = 0; k < nd.nc; ++k) chunks_next[nd.c0 + k] = make_int2(j, nd.start + k * kChunk);
  } else if (f_small[c]) {
    SNode s;
    s.start = ch.x;
    s.end = ch.y;
    s.prec = ch.z;
    s.depth = ch.w;
    small[small_base + o_small[c]] = s;
  } else {
    float lo[3], hi[3];
    seq_box(perm, tri, ch.x, ch.y, lo, hi);
    put_child(out, ch.z, lo, hi, leaf_ref(ch.x, ch.y - ch.x));
  }
  if (c == nch - 1) {
    ctr[C_NB] = o_big[c] + f_big[c];
    ctr[C_NC] = o_chunk[c] + f_chunk[c];
    ctr[C_NS] = o_small[c] + f_small[c];
    ctr[C_NTH] = 0;
  }
}

// Subtrees of at most kSmall triangles: one thread each restates
// build_bvh_recursive (bvh_kernels.cu:667-854), depth first, records
// rec0, rec0 + 1, ... in visit order.
__global__ void k_sah_small(int64_t ns, const SNode* __restrict__ small, const int32_t* __restrict__ rec_off,
                            int32_t rec_base, int leaf_n, int32_t* __restrict__ perm, const SahTri* __restrict__ tri,
                            const float* __restrict__ cz, Node64* __restrict__ out, int32_t* __restrict__ ctr) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const SNode root = small[k];
  int32_t rec = rec_base + rec_off[k];
  if (k == ns - 1) ctr[C_NREC] = rec + (root.end - root.start - 1);
  SNode stack[kSmall + 2];  // pending right children of the ancestors + the two pushed last
  int sp = 0;
  stack[sp++] = root;
  int32_t dmax = 0;
  while (sp > 0) {
    const SNode it = stack[--sp];
    float lo[3], hi[3];
    seq_box(perm, tri, it.start, it.end, lo, hi);
    const int32_t count = it.end - it.start;
    if (count <= leaf_n) {
      put_child(out, it.prec, lo, hi, leaf_ref(it.start, count));
      continue;
    }
    const int32_t self = rec++;
    put_child(out, it.prec, lo, hi, self);
    dmax = it.depth > dmax ? it.depth : dmax;
    const float dims[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    float best_cost = 1e30f;
    int best_axis = -1, best_split = -1;
    for (int ax = 0; ax < 3; ++ax) {
      if (dims[ax] < 1e-6f) continue;
      float clo = 1e30f, chi = -1e30f;
      for (int32_t i = it.start; i < it.end; ++i) {
        const float c = ckey(tri, cz, perm[i], ax);
        clo = fminf(clo, c);
        chi = fmaxf(chi, c);
      }
      if (chi - clo < 1e-6f) continue;
      int cnt[kBins];
      float blo[kBins * 3], bhi[kBins * 3];
      for (int b = 0; b < kBins; ++b) {
        cnt[b] = 0;
        for (int q = 0; q < 3; ++q) {
          blo[b * 3 + q] = 1e30f;
          bhi[b * 3 + q] = -1e30f;
        }
      }
      const float scale = bin_scale(clo, chi);
      for (int32_t i = it.start; i < it.end; ++i) {
        const int32_t t = perm[i];
        const SahTri s = tri[t];
        const int b = bin_of(ckey(tri, cz, t, ax), clo, scale);
        blo[b * 3 + 0] = fminf(blo[b * 3 + 0], s.lo.x);
        blo[b * 3 + 1] = fminf(blo[b * 3 + 1], s.lo.y);
        blo[b * 3 + 2] = fminf(blo[b * 3 + 2], s.lo.z);
        bhi[b * 3 + 0] = fmaxf(bhi[b * 3 + 0], s.hi.x);
        bhi[b * 3 + 1] = fmaxf(bhi[b * 3 + 1], s.hi.y);
        bhi[b * 3 + 2] = fmaxf(bhi[b * 3 + 2], s.hi.z);
        cnt[b]++;
      }
      axis_sweep(cnt, blo, bhi, ax, best_cost, best_axis, best_split);
    }
    int32_t mid = it.start + count / 2;
    if (best_axis != -1) {
      float clo = 1e30f, chi = -1e30f;
      for (int32_t i = it.start; i < it.end; ++i) {
        const float c = ckey(tri, cz, perm[i], best_axis);
        clo = fminf(clo, c);
        chi = fmaxf(chi, c);
      }
      const float sv = split_value(clo, chi, best_split);
      mid = seq_partition(perm, it.start, it.end, tri, cz, best_axis, sv);
      if (mid == it.start || mid == it.end) {
        seq_nth_element(perm, tri, cz, best_axis, it.start, it.start + count / 2, it.end);
        mid = it.start + count / 2;
      }
    } else {
      seq_nth_element(perm, tri, cz, fallback_axis(dims), it.start, it.start + count / 2, it.end);
      mid = it.start + count / 2;
    }
    SNode r, l;
    r.start = mid;
    r.end = it.end;
    r.prec = self * 2 + 1;
    r.depth = it.depth + 1;
    l.start = it.start;
    l.end = mid;
    l.prec = self * 2;
    l.depth = it.depth + 1;
    stack[sp++] = r;
    stack[sp++] = l;
  }
  atomicMax(&ctr[C_DEPTH], dmax);
}

__global__ __launch_bounds__(kBlk) void k_sah_small_counts(int64_t ns, const SNode* __restrict__ small,
                                                            int32_t* __restrict__ sc) {
  const int64_t k = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (k < ns) sc[k] = small[k].end - small[k].start - 1;
}

// Level 0 (the root over [0, n) with its chunk table) or, for n <= kSmall,
// the single small subtree.
__global__ __launch_bounds__(kBlk) void k_sah_root(int64_t n, LNode* __restrict__ lv, int2* __restrict__ chunks,
                                                    SNode* __restrict__ small) {
  const int64_t c = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  const int32_t nc = (int32_t)((n + kChunk - 1) / kChunk);
  if (n <= kSmall) {
    if (c == 0) {
      SNode r;
      r.start = 0;
      r.end = (int32_t)n;
      r.prec = -1;
      r.depth = 0;
      small[0] = r;
    }
    return;
  }
  if (c == 0) {
    LNode r;
    r.start = 0;
    r.end = (int32_t)n;
    r.rec = 0;
    r.prec = -1;
    r.c0 = 0;
    r.nc = nc;
    r.depth = 0;
    r.pad = 0;
    lv[0] = r;
  }
  if (c < nc) chunks[c] = make_int2(0, (int32_t)c * kChunk);
}

// Root leaf (F <= leaf_n, including F = 0): left child = all triangles, right
// child empty (inverted box, never overlapped).
__global__ void k_sah_root_leaf(int64_t n, const int32_t* __restrict__ perm, const SahTri* __restrict__ tri,
                                Node64* __restrict__ out) {
  if (threadIdx.x != 0) return;
  float lo[3], hi[3];
  seq_box(perm, tri, 0, (int32_t)n, lo, hi);
  const float elo[3] = {INFINITY, INFINITY, INFINITY}, ehi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (n > 0) put_child(out, 0, lo, hi, leaf_ref(0, (int32_t)n));
  else put_child(out, 0, elo, ehi, leaf_ref(0, 1));
  put_child(out, 1, elo, ehi, leaf_ref(0, 1));
}

__global__ __launch_bounds__(kBlk) void k_sah_tris(const float* __restrict__ V, const int32_t* __restrict__ F,
                                                    int64_t n, const int32_t* __restrict__ perm,
                                                    Tri48* __restrict__ tris) {
  const int64_t p = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (p >= n) return;
  const int32_t t = perm[p];
  const int64_t i0 = F[(int64_t)t * 3], i1 = F[(int64_t)t * 3 + 1], i2 = F[(int64_t)t * 3 + 2];
  Tri48 o;
  o.p = make_float4(V[i0 * 3], V[i0 * 3 + 1], V[i0 * 3 + 2], V[i1 * 3]);
  o.q = make_float4(V[i1 * 3 + 1], V[i1 * 3 + 2], V[i2 * 3], V[i2 * 3 + 1]);
  o.r = make_float4(V[i2 * 3 + 2], __int_as_float(t), 0.0f, 0.0f);
  tris[p] = o;
}

__global__ void k_sah_header(BvhHeader* __restrict__ h, int64_t n, int64_t cap, int leaf_n, int32_t depth_a,
                             int32_t nrec_a, const int32_t* __restrict__ ctr) {
  if (threadIdx.x != 0) return;
  BvhHeader H;
  for (int k = 0; k < 40; ++k) H.pad[k] = 0;
  H.magic = kBvhMagic;
  H.n_tris = (int32_t)n;
  const int32_t nrec = ctr[C_NREC] > nrec_a ? ctr[C_NREC] : nrec_a;
  H.n_nodes = nrec > 1 ? nrec : 1;
  H.n_node_cap = (int32_t)cap;
  H.nodes_off = 256;
  H.tris_off = 256 + 64 * cap;
  float cm = 0.0f;
  for (int k = 0; k < 3; ++k) {
    H.scene_lo[k] = n > 0 ? ord2f((uint32_t)ctr[C_RED + k]) : 0.0f;
    H.scene_hi[k] = n > 0 ? ord2f((uint32_t)ctr[C_RED + 3 + k]) : 0.0f;
    cm = fmaxf(cm, fmaxf(fabsf(H.scene_lo[k]), fabsf(H.scene_hi[k])));
  }
  H.coord_max = cm;
  H.tri_pad_max = 0.0f;
  H.leaf_max = leaf_n;
  H.depth = ctr[C_DEPTH] > depth_a ? ctr[C_DEPTH] : depth_a;
  H.boxes_off = 0;
  H.obbs_off = 0;
  H.tobbs_off = 0;
  H.pad[0] = kBlobSah;
  *h = H;
}

inline int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

struct SahWs {
  SahTri* tri;
  float* cz;
  int32_t* perm;
  int32_t* pos_lf;
  int32_t* pos_rt;
  LNode* lv[2];
  uint32_t* acc;
  uint32_t* bins;
  int32_t* axis;
  float* split;
  int32_t* mode;
  int32_t* swaps;
  int2* chunks[2];
  int32_t* chunk_true;
  int32_t* chunk_pre;
  int4* child;
  int32_t* f_big;
  int32_t* f_small;
  int32_t* f_chunk;
  int32_t* o_big;
  int32_t* o_small;
  int32_t* o_chunk;
  int32_t* nth_list;
  SNode* small;
  int32_t* small_cnt;
  int32_t* small_off;
  int32_t* ctr;
  void* cub;
  size_t cub_bytes;
  int64_t total;
  int64_t nb_max, nc_max, ns_max;
};

size_t sah_cub_bytes(int64_t n) {
  size_t b = 0;
  (void)a3d::exclusive_sum(nullptr, b, (int32_t*)nullptr, (int32_t*)nullptr, (int)(n > 1 ? n : 2));
  return b;
}

SahWs sah_carve(void* base, int64_t n, int leaf_n) {
  SahWs w;
  char* p = (char*)base;
  int64_t off = 0;
  auto take = [&](int64_t bytes) -> char* {
    char* r = p ? p + off : nullptr;
    off += al256(bytes > 0 ? bytes : 1);
    return r;
  };
  const int64_t nn = n > 0 ? n : 1;
  w.nb_max = nn / (kSmall + 1) + 1;
  w.nc_max = nn / kChunk + w.nb_max + 1;
  w.ns_max = nn / (leaf_n + 1) + 1;
  w.tri = (SahTri*)take(32 * nn);
  w.cz = (float*)take(4 * nn);
  w.perm = (int32_t*)take(4 * nn);
  w.pos_lf = (int32_t*)take(4 * nn);
  w.pos_rt = (int32_t*)take(4 * nn);
  w.lv[0] = (LNode*)take(sizeof(LNode) * w.nb_max);
  w.lv[1] = (LNode*)take(sizeof(LNode) * w.nb_max);
  w.acc = (uint32_t*)take(4 * 12 * w.nb_max);
  w.bins = (uint32_t*)take(4 * (int64_t)kNodeBins * w.nb_max);
  w.axis = (int32_t*)take(4 * w.nb_max);
  w.split = (float*)take(4 * w.nb_max);
  w.mode = (int32_t*)take(4 * w.nb_max);
  w.swaps = (int32_t*)take(4 * w.nb_max);
  w.chunks[0] = (int2*)take(8 * w.nc_max);
  w.chunks[1] = (int2*)take(8 * w.nc_max);
  w.chunk_true = (int32_t*)take(4 * (w.nc_max + 1));
  w.chunk_pre = (int32_t*)take(4 * (w.nc_max + 1));
  const int64_t nch = 2 * w.nb_max;
  w.child = (int4*)take(16 * nch);
  w.f_big = (int32_t*)take(4 * nch);
  w.f_small = (int32_t*)take(4 * nch);
  w.f_chunk = (int32_t*)take(4 * nch);
  w.o_big = (int32_t*)take(4 * nch);
  w.o_small = (int32_t*)take(4 * nch);
  w.o_chunk = (int32_t*)take(4 * nch);
  w.nth_list = (int32_t*)take(4 * w.nb_max);
  w.small = (SNode*)take(sizeof(SNode) * w.ns_max);
  w.small_cnt = (int32_t*)take(4 * w.ns_max);
  w.small_off = (int32_t*)take(4 * w.ns_max);
  w.ctr = (int32_t*)take(4 * C_N);
  int64_t mx = w.nc_max + 1;
  if (nch > mx) mx = nch;
  if (w.ns_max > mx) mx = w.ns_max;
  w.cub_bytes = sah_cub_bytes(mx);
  w.cub = take((int64_t)w.cub_bytes);
  w.total = off;
  return w;
}

inline unsigned gr(int64_t n, int b = kBlk) { return (unsigned)div_up(n > 0 ? n : 1, b); }

}  // namespace
}  // namespace a3d

using namespace a3d;

extern "C" int a3d_sah_sizes(int64_t nF, int leaf_max, int64_t* blob_bytes, int64_t* ws_bytes) {
  if (nF < 0 || nF > (int64_t(1) << 28)) return fail(A3D_ERR_ARG, "a3d_sah_sizes: nF=%lld out of range [0, 2^28]", (long long)nF);
  if (leaf_max < 1 || leaf_max > 8) return fail(A3D_ERR_ARG, "a3d_sah_sizes: leaf_max=%d not in [1, 8]", leaf_max);
  if (!blob_bytes || !ws_bytes) return fail(A3D_ERR_ARG, "a3d_sah_sizes: null output");
  const int64_t cap = nF > 1 ? nF - 1 : 1;
  *blob_bytes = 256 + 64 * cap + 48 * (nF > 0 ? nF : 1);
  *ws_bytes = sah_carve(nullptr, nF, leaf_max).total;
  return A3D_OK;
}

extern "C" int a3d_sah_build(const float* V, int64_t nV, const int32_t* F, int64_t nF, int leaf_max, void* blob,
                             int64_t blob_bytes, void* ws, int64_t ws_bytes, int32_t* depth_out,
                             a3d_stream_t stream) {
  int64_t need_blob = 0, need_ws = 0;
  if (int rc = a3d_sah_sizes(nF, leaf_max, &need_blob, &need_ws)) return rc;
  if (!blob || !ws) return fail(A3D_ERR_ARG, "a3d_sah_build: null blob/workspace");
  if (nF > 0 && (!V || !F || nV <= 0)) return fail(A3D_ERR_ARG, "a3d_sah_build: null or empty mesh arrays");
  if (blob_bytes < need_blob) return fail(A3D_ERR_CAPACITY, "a3d_sah_build: blob %lld < %lld bytes", (long long)blob_bytes, (long long)need_blob);
  if (ws_bytes < need_ws) return fail(A3D_ERR_CAPACITY, "a3d_sah_build: workspace %lld < %lld bytes", (long long)ws_bytes, (long long)need_ws);
  if (((uintptr_t)blob & 255) || ((uintptr_t)ws & 255)) return fail(A3D_ERR_ARG, "a3d_sah_build: blob/workspace must be 256-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = nF;
  const int64_t cap = n > 1 ? n - 1 : 1;
  SahWs w = sah_carve(ws, n, leaf_max);
  char* base = (char*)blob;
  Node64* nodes = (Node64*)(base + 256);
  Tri48* tris = (Tri48*)(base + 256 + 64 * cap);

  hipLaunchKernelGGL(k_sah_init, dim3(1), dim3(64), 0, s, w.ctr);
  A3D_LAUNCH_CHECK();
  if (n > 0) {
    hipLaunchKernelGGL(k_sah_prep, dim3(gr(n) < 2048 ? gr(n) : 2048), dim3(kBlk), 0, s, V, F, n, w.tri, w.cz,
                       w.perm, w.ctr);
    A3D_LAUNCH_CHECK();
  }
  int32_t depth_a = 0, nrec_a = 0;
  if (n <= leaf_max) {
    hipLaunchKernelGGL(k_sah_root_leaf, dim3(1), dim3(64), 0, s, n, w.perm, w.tri, nodes);
    A3D_LAUNCH_CHECK();
    nrec_a = 1;
  } else {
    int64_t ns = 0;
    int32_t rec_base = 0;
    hipLaunchKernelGGL(k_sah_root, dim3(gr(div_up(n, kChunk))), dim3(kBlk), 0, s, n, w.lv[0], w.chunks[0], w.small);
    A3D_LAUNCH_CHECK();
    if (n <= kSmall) {
      ns = 1;
    } else {
      int64_t nb = 1, nc = div_up(n, kChunk);
      int cur = 0;
      int level = 0;
      while (nb > 0) {
        if (level > 4096) return fail(A3D_ERR_INTERNAL, "a3d_sah_build: level loop did not finish");
        const LNode* lv = w.lv[cur];
        const int2* chunks = w.chunks[cur];
        hipLaunchKernelGGL(k_sah_level_init, dim3(gr(nb * kNodeBins)), dim3(kBlk), 0, s, nb, w.acc, w.bins);
        A3D_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sah_stats, dim3((unsigned)nc), dim3(kBlk), 0, s, chunks, lv, w.perm, w.tri, w.cz, w.acc);
        A3D_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sah_bins, dim3((unsigned)nc), dim3(kBlk), 0, s, chunks, lv, w.perm, w.tri, w.cz, w.acc,
                           leaf_max, w.bins);
        A3D_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sah_split, dim3(gr(nb)), dim3(kBlk), 0, s, nb, lv, w.acc, w.bins, nodes, w.axis, w.split,
                           w.mode);
        A3D_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sah_count, dim3((unsigned)nc), dim3(kBlk), 0, s, nc, chunks, lv, w.perm, w.tri, w.cz,
                           w.axis, w.split, w.mode, leaf_max, w.chunk_true);
        A3D_LAUNCH_CHECK();
        size_t cb = w.cub_bytes;
        A3D_HIP_CHECK(a3d::exclusive_sum(w.cub, cb, w.chunk_true, w.chunk_pre, (int)(nc + 1), s));
        hipLaunchKernelGGL(k_sah_rank, dim3((unsigned)nc), dim3(kBlk), 0, s, chunks, lv, w.perm, w.tri, w.cz, w.axis,
                           w.split, w.mode, w.chunk_pre, leaf_max, w.pos_lf, w.pos_rt, w.swaps);
        A3D_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sah_swap, dim3((unsigned)nc), dim3(kBlk), 0, s, chunks, lv, w.mode, w.chunk_pre, leaf_max,
                           w.pos_lf, w.pos_rt, w.swaps, w.perm);
        A3D_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sah_children, dim3(gr(nb)), dim3(kBlk), 0, s, nb, lv, w.axis, w.mode, w.chunk_pre,
                           leaf_max, w.child, w.f_big, w.f_small, w.f_chunk, w.nth_list, w.ctr);
        A3D_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sah_nth, dim3(gr(nb, 64)), dim3(64), 0, s, lv, w.nth_list, w.axis, w.ctr, w.perm, w.tri,
                           w.cz);
        A3D_LAUNCH_CHECK();
        const int64_t nch = 2 * nb;
        cb = w.cub_bytes;
        A3D_HIP_CHECK(a3d::exclusive_sum(w.cub, cb, w.f_big, w.o_big, (int)nch, s));
        cb = w.cub_bytes;
        A3D_HIP_CHECK(a3d::exclusive_sum(w.cub, cb, w.f_small, w.o_small, (int)nch, s));
        cb = w.cub_bytes;
        A3D_HIP_CHECK(a3d::exclusive_sum(w.cub, cb, w.f_chunk, w.o_chunk, (int)nch, s));
        if (ns + nch > w.ns_max) return fail(A3D_ERR_INTERNAL, "a3d_sah_build: small-subtree list overflow");
        hipLaunchKernelGGL(k_sah_emit, dim3(gr(nch)), dim3(kBlk), 0, s, nch, w.child, w.f_big, w.f_small, w.f_chunk,
                           w.o_big, w.o_small, w.o_chunk, (int32_t)(rec_base + nb), (int32_t)ns, w.nb_max, w.nc_max,
                           w.lv[cur ^ 1],
                           w.chunks[cur ^ 1], w.small, w.perm, w.tri, nodes, w.ctr);
        A3D_LAUNCH_CHECK();
        int32_t c3[C_ERR + 1];
        A3D_HIP_CHECK(hipMemcpyAsync(c3, w.ctr, sizeof(c3), hipMemcpyDeviceToHost, s));
        A3D_HIP_CHECK(hipStreamSynchronize(s));
        if (c3[C_ERR] || c3[C_NB] > w.nb_max || c3[C_NC] > w.nc_max)
          return fail(A3D_ERR_INTERNAL, "a3d_sah_build: level %d overflows its node or chunk table", level);
        depth_a = level;
        rec_base += (int32_t)nb;
        nb = c3[C_NB];
        nc = c3[C_NC];
        ns += c3[C_NS];
        cur ^= 1;
        ++level;
      }
    }
    nrec_a = rec_base;
    if (ns > 0) {
      hipLaunchKernelGGL(k_sah_small_counts, dim3(gr(ns)), dim3(kBlk), 0, s, ns, w.small, w.small_cnt);
      A3D_LAUNCH_CHECK();
      size_t cb = w.cub_bytes;
      A3D_HIP_CHECK(a3d::exclusive_sum(w.cub, cb, w.small_cnt, w.small_off, (int)ns, s));
      hipLaunchKernelGGL(k_sah_small, dim3(gr(ns, 64)), dim3(64), 0, s, ns, w.small, w.small_off, rec_base, leaf_max,
                         w.perm, w.tri, w.cz, nodes, w.ctr);
      A3D_LAUNCH_CHECK();
    }
  }
  if (n > 0) {
    hipLaunchKernelGGL(k_sah_tris, dim3(gr(n)), dim3(kBlk), 0, s, V, F, n, w.perm, tris);
    A3D_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_sah_header, dim3(1), dim3(64), 0, s, (BvhHeader*)blob, n, cap, leaf_max, depth_a, nrec_a,
                     w.ctr);
  A3D_LAUNCH_CHECK();
  int32_t depth = 0;
  A3D_HIP_CHECK(hipMemcpyAsync(&depth, (const char*)blob + offsetof(BvhHeader, depth), 4, hipMemcpyDeviceToHost, s));
  A3D_HIP_CHECK(hipStreamSynchronize(s));
  if (depth_out) *depth_out = depth;
  if (depth >= kStackTotal)
    return fail(A3D_ERR_CAPACITY, "a3d_sah_build: tree depth %d exceeds the box walks' stack of %d entries", depth,
                kStackTotal);
  return A3D_OK;
}
