// Shared device code for the atom3d_amd HIP library (gfx950).
//
// Geometry predicates follow the reference's FP32 operation order exactly
// (Atom3D atom3d/kernels/bvh_kernels.cu:77-353) and the library is compiled
// with -ffp-contract=off and IEEE division/sqrt, so every predicate rounds the
// same way on the GPU as the CPU oracle does.  That is what lets closest-face
// ids, SAT hit masks and octree voxel sets be compared bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/atom3d_amd.h"

namespace a3d {

// ------------------------------------------------------------------ errors
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define A3D_HIP_CHECK(expr)                                                     \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess)                                                       \
      return ::a3d::fail(A3D_ERR_HIP, "%s failed: %s (%s:%d)", #expr,           \
                         hipGetErrorString(_e), __FILE__, __LINE__);            \
  } while (0)

#define A3D_LAUNCH_CHECK() A3D_HIP_CHECK(hipGetLastError())

// ------------------------------------------------------------------ vectors
struct v3 { float x, y, z; };
__host__ __device__ __forceinline__ v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
__host__ __device__ __forceinline__ float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ v3 cross3(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__host__ __device__ __forceinline__ v3 sub3(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ __forceinline__ v3 add3(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ __forceinline__ v3 mul3(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fmaxf(lo, fminf(hi, x)); }
__device__ __forceinline__ float sgn(float x) { return x >= 0.0f ? 1.0f : -1.0f; }

// ------------------------------------------------------------------ layout
// BVH blob (one device allocation, owned by the caller):
//   [0, 256)            BvhHeader
//   nodes               Node64[n_node_cap]   internal nodes, root = node 0
//   tris                Tri48[n_tris]        triangles in leaf order
//   boxes               TriBox32[n_tris]     padded triangle boxes, leaf order
//   obbs                Obb96[n_node_cap]    per node, both children's oriented boxes
//   tobbs               float4[3 n_tris]     per triangle, its oriented box (leaf order)
// Node64 keeps BOTH children's boxes so one 64-byte fetch decides both
// children (4 x 16-byte loads, no dependent second fetch).
struct BvhHeader {
  int32_t magic;
  int32_t n_tris;
  int32_t n_nodes;      // internal nodes actually used (written by the build)
  int32_t n_node_cap;   // capacity reserved in the blob
  int64_t nodes_off;    // byte offsets from the blob start
  int64_t tris_off;
  float scene_lo[3];    // exact min/max over all referenced vertices
  float scene_hi[3];
  float coord_max;      // max |coordinate| over referenced vertices
  float tri_pad_max;    // max per-triangle distance error bound (see bvh_build)
  int32_t leaf_max;
  int32_t depth;        // max internal-node depth (root = 0), for stack sizing
  int64_t boxes_off;    // per-triangle padded boxes (TriBox32), leaf order
  int64_t obbs_off;     // per-node child oriented boxes (Obb96), 0 = none
  int64_t tobbs_off;    // per-triangle oriented boxes (3 float4 each, leaf order), 0 = none
  int32_t pad[40];
};
static_assert(sizeof(BvhHeader) == 256, "header is 256 bytes");
constexpr int32_t kBvhMagic = 0x41334442;  // 'A3DB'

// Child reference: >= 0 internal node index; < 0 leaf: ~ref = start << 3 | (count-1)
__host__ __device__ __forceinline__ int32_t leaf_ref(int32_t start, int32_t count) {
  return ~((start << 3) | (count - 1));
}
__device__ __forceinline__ int32_t leaf_start(int32_t ref) { return (~ref) >> 3; }
__device__ __forceinline__ int32_t leaf_count(int32_t ref) { return ((~ref) & 7) + 1; }

struct Node64 {
  float4 a;  // lmin.x lmin.y lmin.z lmax.x
  float4 b;  // lmax.y lmax.z rmin.x rmin.y
  float4 c;  // rmin.z rmax.x rmax.y rmax.z
  int4 d;    // left ref, right ref, 0, 0
};
static_assert(sizeof(Node64) == 64, "node is 64 bytes");

// Triangle in leaf order: a, b, c, the original face id (bit-cast) and the
// triangle's distance-error pad (see bvh_build.hip, tri_pad).
struct Tri48 {
  float4 p;  // a.x a.y a.z b.x
  float4 q;  // b.y b.z c.x c.y
  float4 r;  // c.z id pad 0
};
static_assert(sizeof(Tri48) == 48, "triangle is 48 bytes");

// Triangle bounding box widened by the triangle's distance-error pad, in leaf
// order: lo = min(a, b, c) - pad, hi = max(a, b, c) + pad (w unused).  The
// walks bound-test a triangle from these 32 bytes without its vertices.
struct TriBox32 {
  float4 lo;
  float4 hi;
};
static_assert(sizeof(TriBox32) == 32, "triangle box is 32 bytes");

// Oriented boxes of a node's two children, in each child's own frame: n =
// the normalised sum of its triangles' area vectors, t1 a unit vector
// orthogonal to n, t2 = obb_t2(n, t1); every triangle of the child lies in
// {x : lo_a <= a . x <= hi_a} for a = n, t1, t2, each range widened by the
// largest distance-error pad of those triangles.  The distance from p to that
// box (the three per-axis gaps combined, shrunk by 2^-18 for the axes being
// orthonormal only to float precision) bounds the reference-formula distance
// to each triangle from below, as the padded axis-aligned boxes do, and is
// much tighter for curved or tilted patches.  A child of more than kObbCap
// triangles, or whose normals cancel, has axes 0 and infinite ranges (no
// bound).  Per child 3 float4: (n, lo_n), (t1, lo_1), (hi_n, hi_1, lo_2, hi_2).
struct Obb96 {
  float4 c[6];  // left child c[0..2], right child c[3..5]
};
static_assert(sizeof(Obb96) == 96, "oriented-box record is 96 bytes");
#ifndef A3D_OBB_CAP
#define A3D_OBB_CAP 256
#endif
constexpr int kObbCap = A3D_OBB_CAP;

__host__ __device__ __forceinline__ v3 obb_t2(v3 n, v3 t1) { return cross3(n, t1); }

// Lower bound on the squared distance from p to a child's oriented box; eq =
// 4 ulp of |p|_1 covers the FP32 error of each a . p (|a_i| <= 1).
__device__ __forceinline__ float obb_d2(v3 p, float4 u, float4 v, float4 w, float eq) {
  const v3 n = mk(u.x, u.y, u.z), t1 = mk(v.x, v.y, v.z), t2 = obb_t2(n, t1);
  // fused multiply-adds: each dot product errs by at most 2 ulp of |p|_1
  // (unit axes), within eq; the squared sum errs less than the 2^-18 margin
  const float pn = fmaf(n.z, p.z, fmaf(n.y, p.y, n.x * p.x));
  const float p1 = fmaf(t1.z, p.z, fmaf(t1.y, p.y, t1.x * p.x));
  const float p2 = fmaf(t2.z, p.z, fmaf(t2.y, p.y, t2.x * p.x));
  // max(max(lo - a.p, a.p - hi), eq) - eq == max(max(lo - a.p, a.p - hi) - eq, 0)
  const float dn = fmaxf(fmaxf(u.w - pn, pn - w.x), eq) - eq;
  const float d1 = fmaxf(fmaxf(v.w - p1, p1 - w.y), eq) - eq;
  const float d2 = fmaxf(fmaxf(w.z - p2, p2 - w.w), eq) - eq;
  return fmaf(d2, d2, fmaf(d1, d1, dn * dn)) * 0.999996185302734375f;  // x (1 - 2^-18)
}
__device__ __forceinline__ float obb_eq(v3 p) {
  return 4.0f * 5.9604645e-08f * (fabsf(p.x) + fabsf(p.y) + fabsf(p.z));
}

__device__ __forceinline__ void load_tri(const Tri48* __restrict__ t, int32_t i, v3& a, v3& b, v3& c,
                                         int32_t& id) {
  float4 p = t[i].p, q = t[i].q, r = t[i].r;
  a = mk(p.x, p.y, p.z);
  b = mk(p.w, q.x, q.y);
  c = mk(q.z, q.w, r.x);
  id = __float_as_int(r.y);
}

__device__ __forceinline__ float tri_pad_of(const Tri48* __restrict__ t, int32_t i) { return t[i].r.z; }

__host__ __device__ __forceinline__ TriBox32 tri_box_of(v3 a, v3 b, v3 c, float pad) {
  TriBox32 t;
  t.lo = make_float4(fminf(fminf(a.x, b.x), c.x) - pad, fminf(fminf(a.y, b.y), c.y) - pad,
                     fminf(fminf(a.z, b.z), c.z) - pad, 0.0f);
  t.hi = make_float4(fmaxf(fmaxf(a.x, b.x), c.x) + pad, fmaxf(fmaxf(a.y, b.y), c.y) + pad,
                     fmaxf(fmaxf(a.z, b.z), c.z) + pad, 0.0f);
  return t;
}

__host__ __device__ __forceinline__ const BvhHeader* header(const void* blob) {
  return reinterpret_cast<const BvhHeader*>(blob);
}

// ------------------------------------------------------------------ predicates
// Reference Triangle::distance_sq (bvh_kernels.cu:101-135).
__device__ __forceinline__ float tri_dist2(v3 a, v3 b, v3 c, v3 pos) {
  v3 v21 = sub3(b, a), p1 = sub3(pos, a);
  v3 v32 = sub3(c, b), p2 = sub3(pos, b);
  v3 v13 = sub3(a, c), p3 = sub3(pos, c);
  v3 nor = cross3(v21, v13);
  float nor_sq = dot3(nor, nor);
  bool degen = nor_sq < 1e-12f;
  float st = 0.0f;
  if (!degen) {
    st = sgn(dot3(cross3(v21, nor), p1)) + sgn(dot3(cross3(v32, nor), p2)) + sgn(dot3(cross3(v13, nor), p3));
  }
  if (degen || st < 2.0f) {
    float d, e1, e2, e3;
    v3 q;
    d = clampf(dot3(v21, p1) / fmaxf(dot3(v21, v21), 1e-12f), 0.0f, 1.0f);
    q = sub3(mul3(v21, d), p1); e1 = dot3(q, q);
    d = clampf(dot3(v32, p2) / fmaxf(dot3(v32, v32), 1e-12f), 0.0f, 1.0f);
    q = sub3(mul3(v32, d), p2); e2 = dot3(q, q);
    d = clampf(dot3(v13, p3) / fmaxf(dot3(v13, v13), 1e-12f), 0.0f, 1.0f);
    q = sub3(mul3(v13, d), p3); e3 = dot3(q, q);
    return fminf(e1, fminf(e2, e3));
  }
  float d = dot3(nor, p1);
  return d * d / fmaxf(nor_sq, 1e-12f);
}

// Reference Triangle::closest_point (bvh_kernels.cu:138-182).
__device__ __forceinline__ v3 tri_closest(v3 a, v3 b, v3 c, v3 pos) {
  v3 v21 = sub3(b, a), p1 = sub3(pos, a);
  v3 v32 = sub3(c, b), p2 = sub3(pos, b);
  v3 v13 = sub3(a, c), p3 = sub3(pos, c);
  v3 nor = cross3(v21, v13);
  float nor_sq = dot3(nor, nor);
  bool degen = nor_sq < 1e-12f;
  float st = 0.0f;
  if (!degen) {
    st = sgn(dot3(cross3(v21, nor), p1)) + sgn(dot3(cross3(v32, nor), p2)) + sgn(dot3(cross3(v13, nor), p3));
  }
  if (degen || st < 2.0f) {
    float d;
    d = clampf(dot3(v21, p1) / fmaxf(dot3(v21, v21), 1e-12f), 0.0f, 1.0f);
    v3 c1 = add3(a, mul3(v21, d));
    d = clampf(dot3(v32, p2) / fmaxf(dot3(v32, v32), 1e-12f), 0.0f, 1.0f);
    v3 c2 = add3(b, mul3(v32, d));
    d = clampf(dot3(v13, p3) / fmaxf(dot3(v13, v13), 1e-12f), 0.0f, 1.0f);
    v3 c3 = add3(c, mul3(v13, d));
    v3 q1 = sub3(c1, pos), q2 = sub3(c2, pos), q3 = sub3(c3, pos);
    float d1 = dot3(q1, q1), d2 = dot3(q2, q2), d3 = dot3(q3, q3);
    if (d1 < d2 && d1 < d3) return c1;
    if (d2 < d3) return c2;
    return c3;
  }
  float d = dot3(nor, p1);
  v3 proj = mul3(nor, d / fmaxf(nor_sq, 1e-12f));
  return sub3(pos, proj);
}

// Reference Triangle::barycentric (bvh_kernels.cu:205-222).
__device__ __forceinline__ v3 tri_bary(v3 a, v3 b, v3 c, v3 p) {
  v3 v0 = sub3(b, a), v1 = sub3(c, a), v2 = sub3(p, a);
  float d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1);
  float d20 = dot3(v2, v0), d21 = dot3(v2, v1);
  float denom = d00 * d11 - d01 * d01;
  float v = (d11 * d20 - d01 * d21) / fmaxf(denom, 1e-10f);
  float w = (d00 * d21 - d01 * d20) / fmaxf(denom, 1e-10f);
  float u = 1.0f - v - w;
  return mk(u, v, w);
}

// Reference point_aabb_dist_sq (bvh_kernels.cu:248-253).
__device__ __forceinline__ float box_d2(v3 p, float lx, float ly, float lz, float hx, float hy, float hz) {
  float dx = fmaxf(fmaxf(lx - p.x, 0.0f), p.x - hx);
  float dy = fmaxf(fmaxf(ly - p.y, 0.0f), p.y - hy);
  float dz = fmaxf(fmaxf(lz - p.z, 0.0f), p.z - hz);
  // a pruning bound, not a result: fused multiply-adds (less rounding than
  // the separate products and sums the slack of prune_thr is sized for)
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// Lower bound used to skip exact triangle tests: squared distance from p to
// the triangle's box widened by its error pad.  Whenever this exceeds a
// lane's pruning threshold the reference formula cannot rank the triangle at
// or below the lane's best (same argument as for node boxes).
__device__ __forceinline__ float tri_box_d2(v3 a, v3 b, v3 c, float pad, v3 p) {
  return box_d2(p, fminf(fminf(a.x, b.x), c.x) - pad, fminf(fminf(a.y, b.y), c.y) - pad,
                fminf(fminf(a.z, b.z), c.z) - pad, fmaxf(fmaxf(a.x, b.x), c.x) + pad,
                fmaxf(fmaxf(a.y, b.y), c.y) + pad, fmaxf(fmaxf(a.z, b.z), c.z) + pad);
}

// Reference aabb_overlap (bvh_kernels.cu:241-245), inclusive.
__device__ __forceinline__ bool box_overlap(float lx, float ly, float lz, float hx, float hy, float hz, v3 qlo,
                                            v3 qhi) {
  return (lx <= qhi.x && hx >= qlo.x) && (ly <= qhi.y && hy >= qlo.y) && (lz <= qhi.z && hz >= qlo.z);
}

// Reference triangle_aabb_sat (bvh_kernels.cu:285-353), eps 1e-6 on all 13 axes.
__device__ __forceinline__ bool sat_axis(v3 v0, v3 v1, v3 v2, v3 h, v3 axis) {
  const float eps = 1e-6f;
  float p0 = dot3(v0, axis), p1 = dot3(v1, axis), p2 = dot3(v2, axis);
  float mn = fminf(fminf(p0, p1), p2), mx = fmaxf(fmaxf(p0, p1), p2);
  float rad = h.x * fabsf(axis.x) + h.y * fabsf(axis.y) + h.z * fabsf(axis.z) + eps;
  return !(mn > rad || mx < -rad);
}

__device__ __forceinline__ bool tri_aabb_sat(v3 ta, v3 tb, v3 tc, v3 lo, v3 hi) {
  const float eps = 1e-6f;
  v3 ctr = mk((lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f);
  v3 h = mk((hi.x - lo.x) * 0.5f, (hi.y - lo.y) * 0.5f, (hi.z - lo.z) * 0.5f);
  v3 v0 = sub3(ta, ctr), v1 = sub3(tb, ctr), v2 = sub3(tc, ctr);
  float mn, mx;
  mn = fminf(fminf(v0.x, v1.x), v2.x); mx = fmaxf(fmaxf(v0.x, v1.x), v2.x);
  if (mn > h.x + eps || mx < -h.x - eps) return false;
  mn = fminf(fminf(v0.y, v1.y), v2.y); mx = fmaxf(fmaxf(v0.y, v1.y), v2.y);
  if (mn > h.y + eps || mx < -h.y - eps) return false;
  mn = fminf(fminf(v0.z, v1.z), v2.z); mx = fmaxf(fmaxf(v0.z, v1.z), v2.z);
  if (mn > h.z + eps || mx < -h.z - eps) return false;
  v3 e0 = sub3(v1, v0), e1 = sub3(v2, v1), e2 = sub3(v0, v2);
  if (!sat_axis(v0, v1, v2, h, mk(0.0f, e0.z, -e0.y))) return false;
  if (!sat_axis(v0, v1, v2, h, mk(-e0.z, 0.0f, e0.x))) return false;
  if (!sat_axis(v0, v1, v2, h, mk(e0.y, -e0.x, 0.0f))) return false;
  if (!sat_axis(v0, v1, v2, h, mk(0.0f, e1.z, -e1.y))) return false;
  if (!sat_axis(v0, v1, v2, h, mk(-e1.z, 0.0f, e1.x))) return false;
  if (!sat_axis(v0, v1, v2, h, mk(e1.y, -e1.x, 0.0f))) return false;
  if (!sat_axis(v0, v1, v2, h, mk(0.0f, e2.z, -e2.y))) return false;
  if (!sat_axis(v0, v1, v2, h, mk(-e2.z, 0.0f, e2.x))) return false;
  if (!sat_axis(v0, v1, v2, h, mk(e2.y, -e2.x, 0.0f))) return false;
  v3 n = cross3(e0, sub3(v2, v0));
  float d = -dot3(n, v0);
  float r = h.x * fabsf(n.x) + h.y * fabsf(n.y) + h.z * fabsf(n.z) + eps;
  if (fabsf(d) > r) return false;
  return true;
}

// ------------------------------------------------------------------ traversal
constexpr int kStackTotal = 112;  // >= 97: every tree level adds >= 1 bit of Morton prefix
// Deepest internal node (root = 0) every distance/ray walk's stack holds: a
// walk keeps at most one deferred child per internal level above it, so a
// tree of depth D needs D + 1 entries; the wave walk has 100 (udf.hip).
// a3d_bvh_build fails on a deeper tree instead of letting a walk drop work.
constexpr int kMaxWalkDepth = 99;

__device__ __forceinline__ const Node64* nodes_of(const void* blob) {
  return reinterpret_cast<const Node64*>((const char*)blob + header(blob)->nodes_off);
}
__device__ __forceinline__ const Tri48* tris_of(const void* blob) {
  return reinterpret_cast<const Tri48*>((const char*)blob + header(blob)->tris_off);
}
__device__ __forceinline__ const TriBox32* boxes_of(const void* blob) {
  return reinterpret_cast<const TriBox32*>((const char*)blob + header(blob)->boxes_off);
}
// Per-triangle oriented box, same 3-float4 layout as one Obb96 child: axes
// n = the unit normal, t1 = the unit first edge, t2 = obb_t2(n, t1); ranges
// of the vertices widened by the triangle's pad.  Degenerate triangles get
// zero axes and infinite ranges (no bound).
__device__ __forceinline__ const float4* tobbs_of(const void* blob) {
  const int64_t o = header(blob)->tobbs_off;
  return o ? reinterpret_cast<const float4*>((const char*)blob + o) : nullptr;
}
__device__ __forceinline__ const Obb96* obbs_of(const void* blob) {
  const int64_t o = header(blob)->obbs_off;
  return o ? reinterpret_cast<const Obb96*>((const char*)blob + o) : nullptr;
}

// Squared-distance threshold above which a box cannot hold a face the
// reference formula would rank at or below `best` (see the UDF kernels).
// The hardware square root (within 1 ulp, i.e. a factor 1 - 2^-23 at worst)
// is enough: the (1 + 2^-20) widening covers its error.
__device__ __forceinline__ float prune_thr(float best, float slack) {
  float t = (__builtin_amdgcn_sqrtf(best) + slack) * 1.00000095367431640625f;  // x (1 + 2^-20)
  return t * t;
}

// Absolute error slack of a query: 16 ulp of the largest magnitude entering a
// coordinate difference.
__device__ __forceinline__ float query_slack(v3 p, float coord_max) {
  return 16.0f * 5.9604645e-08f * (fabsf(p.x) + fabsf(p.y) + fabsf(p.z) + 3.0f * coord_max);
}

// ------------------------------------------------------------------ launch helpers
// Blocks are handed to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Remap so each XCD walks runs of kXcdChunk consecutive logical blocks
// (neighbouring queries share its L2) and the runs of all XCDs interleave
// over the whole index range, so query regions of uneven cost spread evenly
// over the XCDs.  (One contiguous range per XCD left XCDs idle for up to a
// third of a walk on clustered workloads.)  A trailing partial round of
// chunks maps to itself.
constexpr uint32_t kXcdChunk = 64;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t nx = 8, S = nx * kXcdChunk, full = nb / S * S;
  if (b >= full) return b;
  const uint32_t x = b % nx, k = b / nx;
  return ((k / kXcdChunk) * nx + x) * kXcdChunk + (k % kXcdChunk);
}

inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ------------------------------------------------------------------ lane refill
// Independent per-lane walks (one ray per lane) vary in length by
// orders of magnitude; a wavefront that walked 64 of them side by side would
// run as long as its longest one with most lanes idle.  Instead each
// wavefront of a grid sized to the resident wavefronts works through a list
// of items, and its idle lanes take the next items while the others walk on.
// Wavefront wv of W takes the blocks of 64 consecutive items b = wv, wv + W,
// ...; its job j is item ((j >> 6) * W + wv) * 64 + (j & 63).  (Blocks in
// runs per XCD, as xcd_block maps them, made random rays 7 % slower.)
struct LaneJobs {
  int64_t W, wv, njobs, next;  // wave-uniform
};

__device__ __forceinline__ LaneJobs lane_jobs(int64_t K, int block_threads) {
  LaneJobs J;
  const int wpb = block_threads >> 6;
  J.W = (int64_t)gridDim.x * wpb;
  J.wv = (int64_t)blockIdx.x * wpb + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t nb = (K + 63) >> 6;
  J.njobs = J.wv < nb ? ((nb - 1 - J.wv) / J.W + 1) * 64 : 0;
  J.next = 0;
  return J;
}

// Hands the next jobs to the lanes of `idle` in lane order (call with the
// whole wavefront; the lanes outside `idle` get -1 and keep their item).
// Returns this lane's new item, or -1 (no job left, or a job past K in the
// last block).
__device__ __forceinline__ int64_t lane_take(LaneJobs& J, uint64_t idle, int64_t K) {
  const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
  const int64_t j = J.next + rank;
  const bool mine = (idle >> (threadIdx.x & 63)) & 1ull;
  J.next += __popcll(idle);
  if (!mine || j >= J.njobs) return -1;
  const int64_t item = ((j >> 6) * J.W + J.wv) * 64 + (j & 63);
  return item < K ? item : -1;
}

// Lanes refill once at least this many are idle (or all of them are).
constexpr int kRefillIdle = 16;

// Grid for a refill kernel: at most the blocks resident at once on the
// device (more would start late and end the launch late), at least enough
// for one block of 64 items per wavefront.  cache: the call site's
// per-device record of resident blocks (64 ints, zero-initialised).
template <typename Kern>
inline unsigned refill_grid(Kern kern, int block_threads, int64_t K, int* cache) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (!cache[dev]) {
    int n = 0, c = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, block_threads, 0) != hipSuccess || n < 1) n = 1;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c < 1) c = 1;
    cache[dev] = n * c;
  }
  const int64_t want = div_up(K > 0 ? K : 1, block_threads);
  return (unsigned)(want < cache[dev] ? want : cache[dev]);
}

// ------------------------------------------------------------------ leaf tests dealt over the wavefront
// In a per-lane ray walk a lane that reaches a leaf tests its (up to 8)
// triangles one after another while the other lanes wait; with 64 lanes
// some lane is at a leaf in most steps, so the wavefront would spend most
// of its time in leaf loops with a few lanes active.  Instead the (lane,
// triangle) pairs of every lane at a leaf are dealt out one per lane, in
// rounds of 64, and tested side by side; each result goes back to the lane
// that owns the pair.
__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// This lane's leaf (n triangles from start; n = 0: none) and the exclusive
// prefix of n over the lanes; total pairs (wave-uniform).
struct LeafDeal {
  int32_t start, n, excl, total;
};

__device__ __forceinline__ LeafDeal leaf_deal(bool at_leaf, int32_t ref, int lane) {
  LeafDeal d;
  d.n = at_leaf ? leaf_count(ref) : 0;
  d.start = at_leaf ? leaf_start(ref) : 0;
  int incl = d.n;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  d.excl = incl - d.n;
  d.total = __shfl(incl, 63);
  return d;
}

// Pair base + lane of a round: owner lane L and triangle (valid = the pair
// exists).  own: the wavefront's 64-entry owner table.  Ends with the table
// read; the caller's next round starts with a fence.
__device__ __forceinline__ void leaf_pair(const LeafDeal& d, int base, uint8_t* own, int lane, bool& valid, int& L,
                                          int32_t& tri) {
  const int lo = d.excl > base ? d.excl : base;
  const int hi = d.excl + d.n < base + 64 ? d.excl + d.n : base + 64;
  for (int k = lo; k < hi; ++k) own[k - base] = (uint8_t)lane;
  wave_fence();
  const int j = base + lane;
  valid = j < d.total;
  L = valid ? (int)own[lane] : lane;
  tri = __shfl(d.start, L) + (j - __shfl(d.excl, L));
}

// Lanes [0, k) as a mask, k in [0, 64].
__device__ __forceinline__ uint64_t lanes_below(int k) { return k >= 64 ? ~0ull : ((1ull << k) - 1ull); }

// This lane's pairs among the lanes of round `base`, as a lane mask.
__device__ __forceinline__ uint64_t own_pairs(const LeafDeal& d, int base) {
  const int lo = d.excl - base > 0 ? d.excl - base : 0;
  const int hi = d.excl + d.n - base < 64 ? d.excl + d.n - base : 64;
  return (d.n > 0 && lo < hi) ? (lanes_below(hi) & ~lanes_below(lo)) : 0ull;
}

}  // namespace a3d
