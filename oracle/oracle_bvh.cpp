// Oracle — CPU restatement of the reference's BVH geometry algorithms.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
// atom3d_amd/csrc: it is linked by tests/, by __graft_entry__.smoke() and by
// bench.py's cpu_baseline leg, never by the product path.  Everything here
// runs on host memory with plain C++ (g++ -O3 -ffp-contract=off, IEEE float
// division and sqrt) so that every FP32 operation is rounded exactly once in
// the order the reference writes it.
//
// Restated from (Atom3D, atom3d/kernels/bvh_kernels.cu):
//   Triangle::distance_sq / closest_point / barycentric   :101-135, 138-182, 205-222
//   aabb_overlap, point_aabb_dist_sq                       :241-253
//   triangle_aabb_sat                                      :285-353
//   bvh_udf_kernel (stack walk)                            :363-463
//   bvh_aabb_intersect_kernel (overlap walk + SAT)         :555-610
//   SAHBin / build_bvh_recursive / build_bvh_cuda          :616-925
// plus the octree level loop of atom3d/grid/octree_indexer.py:426-509 and
// cube_aabb_level :302-337.
//
// C++ (not C) because the reference's SAH build splits with std::partition and
// std::nth_element; calling the same libstdc++ algorithms reproduces the
// reference tree layout byte for byte.
//
// Pinning (see tests/test_oracle_bvh.py): the triangle distance is checked
// against an independent FP64 Voronoi-region closest point (Ericson, Real-Time
// Collision Detection §5.1.5) and against closed forms (plane / edge / vertex
// regions of hand-made triangles); the SAH tree is checked by brute force
// (traversal result == all-faces scan) and structural invariants; the SAT is
// checked against exact clipping on random boxes and known touching cases.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

struct f3 { float x, y, z; };
static inline f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross3(f3 a, f3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline f3 sub3(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 add3(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 mul3(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline float clampf(float x, float lo, float hi) { return fmaxf(lo, fminf(hi, x)); }
static inline float signf_(float x) { return x >= 0.0f ? 1.0f : -1.0f; }

// bvh_kernels.cu:101-135
static float tri_dist2(f3 a, f3 b, f3 c, f3 pos) {
  f3 v21 = sub3(b, a), p1 = sub3(pos, a);
  f3 v32 = sub3(c, b), p2 = sub3(pos, b);
  f3 v13 = sub3(a, c), p3 = sub3(pos, c);
  f3 nor = cross3(v21, v13);
  float nor_sq = dot3(nor, nor);
  bool degen = nor_sq < 1e-12f;
  float st = 0.0f;
  if (!degen) {
    st = signf_(dot3(cross3(v21, nor), p1)) + signf_(dot3(cross3(v32, nor), p2)) +
         signf_(dot3(cross3(v13, nor), p3));
  }
  if (degen || st < 2.0f) {
    auto ed = [](f3 v, f3 p) {
      float d = dot3(v, p) / fmaxf(dot3(v, v), 1e-12f);
      d = clampf(d, 0.0f, 1.0f);
      f3 q = sub3(mul3(v, d), p);
      return dot3(q, q);
    };
    return fminf(ed(v21, p1), fminf(ed(v32, p2), ed(v13, p3)));
  }
  float d = dot3(nor, p1);
  return d * d / fmaxf(nor_sq, 1e-12f);
}

// bvh_kernels.cu:138-182
static f3 tri_closest(f3 a, f3 b, f3 c, f3 pos) {
  f3 v21 = sub3(b, a), p1 = sub3(pos, a);
  f3 v32 = sub3(c, b), p2 = sub3(pos, b);
  f3 v13 = sub3(a, c), p3 = sub3(pos, c);
  f3 nor = cross3(v21, v13);
  float nor_sq = dot3(nor, nor);
  bool degen = nor_sq < 1e-12f;
  float st = 0.0f;
  if (!degen) {
    st = signf_(dot3(cross3(v21, nor), p1)) + signf_(dot3(cross3(v32, nor), p2)) +
         signf_(dot3(cross3(v13, nor), p3));
  }
  if (degen || st < 2.0f) {
    auto ec = [](f3 v, f3 p, f3 o) {
      float d = dot3(v, p) / fmaxf(dot3(v, v), 1e-12f);
      d = clampf(d, 0.0f, 1.0f);
      return add3(o, mul3(v, d));
    };
    f3 c1 = ec(v21, p1, a), c2 = ec(v32, p2, b), c3 = ec(v13, p3, c);
    f3 q1 = sub3(c1, pos), q2 = sub3(c2, pos), q3 = sub3(c3, pos);
    float d1 = dot3(q1, q1), d2 = dot3(q2, q2), d3 = dot3(q3, q3);
    if (d1 < d2 && d1 < d3) return c1;
    if (d2 < d3) return c2;
    return c3;
  }
  float d = dot3(nor, p1);
  f3 proj = mul3(nor, d / fmaxf(nor_sq, 1e-12f));
  return sub3(pos, proj);
}

// bvh_kernels.cu:205-222
static f3 tri_bary(f3 a, f3 b, f3 c, f3 p) {
  f3 v0 = sub3(b, a), v1 = sub3(c, a), v2 = sub3(p, a);
  float d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1);
  float d20 = dot3(v2, v0), d21 = dot3(v2, v1);
  float denom = d00 * d11 - d01 * d01;
  float v = (d11 * d20 - d01 * d21) / fmaxf(denom, 1e-10f);
  float w = (d00 * d21 - d01 * d20) / fmaxf(denom, 1e-10f);
  float u = 1.0f - v - w;
  return mk(u, v, w);
}

// bvh_kernels.cu:241-245 (inclusive overlap)
static inline bool aabb_overlap(const float* bmin, const float* bmax, f3 qmin, f3 qmax) {
  return (bmin[0] <= qmax.x && bmax[0] >= qmin.x) && (bmin[1] <= qmax.y && bmax[1] >= qmin.y) &&
         (bmin[2] <= qmax.z && bmax[2] >= qmin.z);
}

// bvh_kernels.cu:248-253
static inline float point_aabb_d2(f3 p, const float* bmin, const float* bmax) {
  float dx = fmaxf(fmaxf(bmin[0] - p.x, 0.0f), p.x - bmax[0]);
  float dy = fmaxf(fmaxf(bmin[1] - p.y, 0.0f), p.y - bmax[1]);
  float dz = fmaxf(fmaxf(bmin[2] - p.z, 0.0f), p.z - bmax[2]);
  return dx * dx + dy * dy + dz * dz;
}

// bvh_kernels.cu:285-353
static bool tri_aabb_sat(f3 ta, f3 tb, f3 tc, f3 bmin, f3 bmax) {
  const float eps = 1e-6f;
  f3 ctr = mk((bmin.x + bmax.x) * 0.5f, (bmin.y + bmax.y) * 0.5f, (bmin.z + bmax.z) * 0.5f);
  f3 h = mk((bmax.x - bmin.x) * 0.5f, (bmax.y - bmin.y) * 0.5f, (bmax.z - bmin.z) * 0.5f);
  f3 v0 = sub3(ta, ctr), v1 = sub3(tb, ctr), v2 = sub3(tc, ctr);
  float mn, mx;
  mn = fminf(fminf(v0.x, v1.x), v2.x); mx = fmaxf(fmaxf(v0.x, v1.x), v2.x);
  if (mn > h.x + eps || mx < -h.x - eps) return false;
  mn = fminf(fminf(v0.y, v1.y), v2.y); mx = fmaxf(fmaxf(v0.y, v1.y), v2.y);
  if (mn > h.y + eps || mx < -h.y - eps) return false;
  mn = fminf(fminf(v0.z, v1.z), v2.z); mx = fmaxf(fmaxf(v0.z, v1.z), v2.z);
  if (mn > h.z + eps || mx < -h.z - eps) return false;
  f3 e[3] = {sub3(v1, v0), sub3(v2, v1), sub3(v0, v2)};
  for (int k = 0; k < 3; ++k) {
    for (int ax = 0; ax < 3; ++ax) {
      f3 E = e[k];
      f3 axis = ax == 0 ? mk(0.0f, E.z, -E.y) : (ax == 1 ? mk(-E.z, 0.0f, E.x) : mk(E.y, -E.x, 0.0f));
      float p0 = dot3(v0, axis), p1 = dot3(v1, axis), p2 = dot3(v2, axis);
      float mnp = fminf(fminf(p0, p1), p2), mxp = fmaxf(fmaxf(p0, p1), p2);
      float rad = h.x * fabsf(axis.x) + h.y * fabsf(axis.y) + h.z * fabsf(axis.z) + eps;
      if (mnp > rad || mxp < -rad) return false;
    }
  }
  f3 n = cross3(e[0], sub3(v2, v0));
  float d = -dot3(n, v0);
  float r = h.x * fabsf(n.x) + h.y * fabsf(n.y) + h.z * fabsf(n.z) + eps;
  if (fabsf(d) > r) return false;
  return true;
}

static inline bool tri_box_overlap(f3 a, f3 b, f3 c, f3 qmin, f3 qmax) {
  float tmin[3] = {fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z)};
  float tmax[3] = {fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z)};
  return aabb_overlap(tmin, tmax, qmin, qmax);
}

// ---------------------------------------------------------------- SAH build
// bvh_kernels.cu:229-235 node layout: bb_min[3], bb_max[3], left, right, pad.
struct Node { float bmin[3]; float bmax[3]; int left, right, pad; };
static_assert(sizeof(Node) == 36, "node is 9 x 4 bytes");

struct TriInfo { f3 centroid; f3 a, b, c; int original_id; int idx; };

struct Bin {
  float bmin[3], bmax[3]; int count;
  Bin() { bmin[0] = bmin[1] = bmin[2] = 1e30f; bmax[0] = bmax[1] = bmax[2] = -1e30f; count = 0; }
  void grow(const f3& a, const f3& b, const f3& c) {
    bmin[0] = fminf(bmin[0], fminf(fminf(a.x, b.x), c.x));
    bmin[1] = fminf(bmin[1], fminf(fminf(a.y, b.y), c.y));
    bmin[2] = fminf(bmin[2], fminf(fminf(a.z, b.z), c.z));
    bmax[0] = fmaxf(bmax[0], fmaxf(fmaxf(a.x, b.x), c.x));
    bmax[1] = fmaxf(bmax[1], fmaxf(fmaxf(a.y, b.y), c.y));
    bmax[2] = fmaxf(bmax[2], fmaxf(fmaxf(a.z, b.z), c.z));
    count++;
  }
  void grow(const Bin& o) {
    for (int k = 0; k < 3; ++k) { bmin[k] = fminf(bmin[k], o.bmin[k]); bmax[k] = fmaxf(bmax[k], o.bmax[k]); }
    count += o.count;
  }
  float area() const {
    if (count == 0) return 0.0f;
    float dx = fmaxf(0.0f, bmax[0] - bmin[0]);
    float dy = fmaxf(0.0f, bmax[1] - bmin[1]);
    float dz = fmaxf(0.0f, bmax[2] - bmin[2]);
    return 2.0f * (dx * dy + dy * dz + dz * dx);
  }
};

static inline float caxis(const TriInfo& t, int ax) {
  return ax == 0 ? t.centroid.x : (ax == 1 ? t.centroid.y : t.centroid.z);
}

// bvh_kernels.cu:667-854 (the recursion is kept; depth is bounded by the
// median fallback to O(log F) for all but adversarial inputs)
static void build_rec(std::vector<TriInfo>& tris, int start, int end, std::vector<Node>& nodes,
                      int node_idx, int leaf_n, int64_t* stats) {
  Node& node = nodes[node_idx];
  node.bmin[0] = node.bmin[1] = node.bmin[2] = 1e30f;
  node.bmax[0] = node.bmax[1] = node.bmax[2] = -1e30f;
  node.pad = 0;
  for (int i = start; i < end; i++) {
    const TriInfo& t = tris[i];
    node.bmin[0] = fminf(node.bmin[0], fminf(fminf(t.a.x, t.b.x), t.c.x));
    node.bmin[1] = fminf(node.bmin[1], fminf(fminf(t.a.y, t.b.y), t.c.y));
    node.bmin[2] = fminf(node.bmin[2], fminf(fminf(t.a.z, t.b.z), t.c.z));
    node.bmax[0] = fmaxf(node.bmax[0], fmaxf(fmaxf(t.a.x, t.b.x), t.c.x));
    node.bmax[1] = fmaxf(node.bmax[1], fmaxf(fmaxf(t.a.y, t.b.y), t.c.y));
    node.bmax[2] = fmaxf(node.bmax[2], fmaxf(fmaxf(t.a.z, t.b.z), t.c.z));
  }
  int count = end - start;
  if (count <= leaf_n) {
    node.left = -(start + 1);
    node.right = -(end + 1);
    return;
  }
  const int NB = 16;
  int best_axis = -1, best_split = -1;
  float best_cost = 1e30f;
  float dims[3] = {node.bmax[0] - node.bmin[0], node.bmax[1] - node.bmin[1], node.bmax[2] - node.bmin[2]};
  for (int axis = 0; axis < 3; axis++) {
    if (dims[axis] < 1e-6f) continue;
    float lo = 1e30f, hi = -1e30f;
    for (int i = start; i < end; ++i) { float c = caxis(tris[i], axis); lo = fminf(lo, c); hi = fmaxf(hi, c); }
    if (hi - lo < 1e-6f) continue;
    Bin bins[NB];
    float scale = NB / (hi - lo);
    scale *= 0.999f;
    for (int i = start; i < end; i++) {
      float c = caxis(tris[i], axis);
      int b = (int)((c - lo) * scale);
      b = std::max(0, std::min(NB - 1, b));
      bins[b].grow(tris[i].a, tris[i].b, tris[i].c);
    }
    float la[NB - 1], ra[NB - 1];
    int lc[NB - 1], rc[NB - 1];
    Bin cur; int cc = 0;
    for (int i = 0; i < NB - 1; i++) { cur.grow(bins[i]); cc += bins[i].count; la[i] = cur.area(); lc[i] = cc; }
    cur = Bin(); cc = 0;
    for (int i = NB - 1; i > 0; i--) { cur.grow(bins[i]); cc += bins[i].count; ra[i - 1] = cur.area(); rc[i - 1] = cc; }
    for (int i = 0; i < NB - 1; i++) {
      float cost = lc[i] * la[i] + rc[i] * ra[i];
      if (cost < best_cost) { best_cost = cost; best_axis = axis; best_split = i; }
    }
  }
  int mid = start + count / 2;
  if (best_axis != -1) {
    float lo = 1e30f, hi = -1e30f;
    for (int i = start; i < end; ++i) { float c = caxis(tris[i], best_axis); lo = fminf(lo, c); hi = fmaxf(hi, c); }
    float split = lo + (best_split + 1) * (hi - lo) / NB;
    int ax = best_axis;
    auto it = std::partition(tris.begin() + start, tris.begin() + end,
                             [&](const TriInfo& t) { return caxis(t, ax) < split; });
    mid = (int)std::distance(tris.begin(), it);
    if (mid == start || mid == end) {
      if (stats) stats[0]++;
      std::nth_element(tris.begin() + start, tris.begin() + start + count / 2, tris.begin() + end,
                       [ax](const TriInfo& a, const TriInfo& b) { return caxis(a, ax) < caxis(b, ax); });
      mid = start + count / 2;
    }
  } else {
    if (stats) stats[1]++;
    int fa = 0;
    if (dims[1] > dims[0]) fa = 1;
    if (dims[2] > dims[fa]) fa = 2;
    std::nth_element(tris.begin() + start, tris.begin() + start + count / 2, tris.begin() + end,
                     [fa](const TriInfo& a, const TriInfo& b) { return caxis(a, fa) < caxis(b, fa); });
    mid = start + count / 2;
  }
  int left = (int)nodes.size();
  nodes.emplace_back();
  nodes.emplace_back();
  nodes[node_idx].left = left;
  nodes[node_idx].right = left + 1;
  build_rec(tris, start, mid, nodes, left, leaf_n, stats);
  build_rec(tris, mid, end, nodes, left + 1, leaf_n, stats);
}

static inline f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

}  // namespace

extern "C" {

int orc_num_threads(void) { return omp_get_max_threads(); }

// Point-triangle squared distance, closest point and barycentrics for one
// (point, triangle) pair (bvh_kernels.cu:101-222).  tri9 = a,b,c.
void orc_tri_query(const float* tri9, const float* p, float* d2, float* closest, float* uvw) {
  f3 a = ld3(tri9), b = ld3(tri9 + 3), c = ld3(tri9 + 6), q = ld3(p);
  *d2 = tri_dist2(a, b, c, q);
  f3 cp = tri_closest(a, b, c, q);
  f3 w = tri_bary(a, b, c, cp);
  closest[0] = cp.x; closest[1] = cp.y; closest[2] = cp.z;
  uvw[0] = w.x; uvw[1] = w.y; uvw[2] = w.z;
}

int orc_tri_aabb_sat(const float* tri9, const float* bmin, const float* bmax) {
  return tri_aabb_sat(ld3(tri9), ld3(tri9 + 3), ld3(tri9 + 6), ld3(bmin), ld3(bmax)) ? 1 : 0;
}

// Reference SAH build (bvh_kernels.cu:860-925).  nodes_out must hold
// 2*max(F,1) nodes of 9 floats; tris_out holds F rows of 10 floats (a, b, c,
// original id bit-cast).  Returns the node count.  stats (nullable, 2
// entries) counts the median fallbacks: [0] after an empty SAH partition,
// [1] with no axis to bin.
int64_t orc_sah_build(const float* V, int64_t nV, const int32_t* F, int64_t nF, int leaf_n,
                      float* nodes_out, float* tris_out, int64_t* stats) {
  (void)nV;
  std::vector<TriInfo> ti((size_t)nF);
  for (int64_t i = 0; i < nF; i++) {
    int i0 = F[i * 3], i1 = F[i * 3 + 1], i2 = F[i * 3 + 2];
    TriInfo& t = ti[(size_t)i];
    t.a = ld3(V + (int64_t)i0 * 3); t.b = ld3(V + (int64_t)i1 * 3); t.c = ld3(V + (int64_t)i2 * 3);
    t.original_id = (int)i; t.idx = 0;
    t.centroid = mk((t.a.x + t.b.x + t.c.x) / 3.0f, (t.a.y + t.b.y + t.c.y) / 3.0f,
                    (t.a.z + t.b.z + t.c.z) / 3.0f);
  }
  std::vector<Node> nodes;
  nodes.emplace_back();
  nodes.reserve((size_t)nF * 2 + 1);
  build_rec(ti, 0, (int)nF, nodes, 0, leaf_n, stats);
  std::memcpy(nodes_out, nodes.data(), nodes.size() * sizeof(Node));
  for (int64_t i = 0; i < nF; i++) {
    const TriInfo& t = ti[(size_t)i];
    float* o = tris_out + i * 10;
    o[0] = t.a.x; o[1] = t.a.y; o[2] = t.a.z; o[3] = t.b.x; o[4] = t.b.y; o[5] = t.b.z;
    o[6] = t.c.x; o[7] = t.c.y; o[8] = t.c.z;
    std::memcpy(o + 9, &t.original_id, 4);
  }
  return (int64_t)nodes.size();
}

// Reference UDF traversal (bvh_kernels.cu:363-463) over a tree from
// orc_sah_build.  uvw of a query with no face is left as zeros (the reference
// leaves it uninitialised).  visits/tests (nullable) count node pops that pass
// the prune test and triangle tests, per query.
void orc_bvh_udf(const float* nodes_f, const float* tris, const float* P, int64_t K, float* dist,
                 int32_t* face, float* closest, float* uvw, int32_t* visits, int32_t* tests,
                 int nthreads) {
  const Node* nodes = reinterpret_cast<const Node*>(nodes_f);
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t q = 0; q < K; q++) {
    f3 p = ld3(P + q * 3);
    float best = 1e30f;
    int best_face = -1, best_tri = -1;
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    int nv = 0, nt = 0;
    while (sp > 0) {
      int idx = stack[--sp];
      const Node& n = nodes[idx];
      float nd = point_aabb_d2(p, n.bmin, n.bmax);
      if (nd >= best) continue;
      nv++;
      if (n.left < 0) {
        int s = -n.left - 1, e = -n.right - 1;
        for (int i = s; i < e; i++) {
          const float* t = tris + (int64_t)i * 10;
          float d2 = tri_dist2(ld3(t), ld3(t + 3), ld3(t + 6), p);
          nt++;
          if (d2 < best) { best = d2; std::memcpy(&best_face, t + 9, 4); best_tri = i; }
        }
      } else {
        int l = n.left, r = n.right;
        float d1 = point_aabb_d2(p, nodes[l].bmin, nodes[l].bmax);
        float d2 = point_aabb_d2(p, nodes[r].bmin, nodes[r].bmax);
        if (d1 < d2) {
          if (d2 < best && sp < 128) stack[sp++] = r;
          if (d1 < best && sp < 128) stack[sp++] = l;
        } else {
          if (d1 < best && sp < 128) stack[sp++] = l;
          if (d2 < best && sp < 128) stack[sp++] = r;
        }
      }
    }
    dist[q] = sqrtf(best);
    face[q] = best_face;
    f3 cp = p;
    f3 w = mk(0.0f, 0.0f, 0.0f);
    if (best_tri >= 0) {
      const float* t = tris + (int64_t)best_tri * 10;
      cp = tri_closest(ld3(t), ld3(t + 3), ld3(t + 6), p);
      w = tri_bary(ld3(t), ld3(t + 3), ld3(t + 6), cp);
    }
    closest[q * 3] = cp.x; closest[q * 3 + 1] = cp.y; closest[q * 3 + 2] = cp.z;
    uvw[q * 3] = w.x; uvw[q * 3 + 1] = w.y; uvw[q * 3 + 2] = w.z;
    if (visits) visits[q] = nv;
    if (tests) tests[q] = nt;
  }
}

// Tree-independent closest face: scan every face in index order, keep the
// first strict minimum of the reference's FP32 squared distance, i.e. ties go
// to the lowest face index.  Same outputs as orc_bvh_udf.
void orc_brute_udf(const float* V, const int32_t* F, int64_t nF, const float* P, int64_t K,
                   float* dist, int32_t* face, float* closest, float* uvw, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t q = 0; q < K; q++) {
    f3 p = ld3(P + q * 3);
    float best = 1e30f;
    int64_t bf = -1;
    for (int64_t f = 0; f < nF; f++) {
      f3 a = ld3(V + (int64_t)F[f * 3] * 3), b = ld3(V + (int64_t)F[f * 3 + 1] * 3),
         c = ld3(V + (int64_t)F[f * 3 + 2] * 3);
      float d2 = tri_dist2(a, b, c, p);
      if (d2 < best) { best = d2; bf = f; }
    }
    dist[q] = sqrtf(best);
    face[q] = (int32_t)bf;
    f3 cp = p, w = mk(0.0f, 0.0f, 0.0f);
    if (bf >= 0) {
      f3 a = ld3(V + (int64_t)F[bf * 3] * 3), b = ld3(V + (int64_t)F[bf * 3 + 1] * 3),
         c = ld3(V + (int64_t)F[bf * 3 + 2] * 3);
      cp = tri_closest(a, b, c, p);
      w = tri_bary(a, b, c, cp);
    }
    closest[q * 3] = cp.x; closest[q * 3 + 1] = cp.y; closest[q * 3 + 2] = cp.z;
    uvw[q * 3] = w.x; uvw[q * 3 + 1] = w.y; uvw[q * 3 + 2] = w.z;
  }
}

// FP64 exact point-triangle closest point by Voronoi regions (Ericson, Real-Time
// Collision Detection, §5.1.5).  Independent of the reference's formulation;
// used to state the error of the FP32 results.
static void closest_fp64(const double* a, const double* b, const double* c, const double* p,
                         double* out) {
  double ab[3], ac[3], ap[3];
  for (int k = 0; k < 3; ++k) { ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; ap[k] = p[k] - a[k]; }
  auto dt = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  double d1 = dt(ab, ap), d2 = dt(ac, ap);
  if (d1 <= 0 && d2 <= 0) { for (int k = 0; k < 3; ++k) out[k] = a[k]; return; }
  double bp[3]; for (int k = 0; k < 3; ++k) bp[k] = p[k] - b[k];
  double d3 = dt(ab, bp), d4 = dt(ac, bp);
  if (d3 >= 0 && d4 <= d3) { for (int k = 0; k < 3; ++k) out[k] = b[k]; return; }
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    double den = d1 - d3;
    double v = den != 0 ? d1 / den : 0.0;
    for (int k = 0; k < 3; ++k) out[k] = a[k] + v * ab[k];
    return;
  }
  double cp[3]; for (int k = 0; k < 3; ++k) cp[k] = p[k] - c[k];
  double d5 = dt(ab, cp), d6 = dt(ac, cp);
  if (d6 >= 0 && d5 <= d6) { for (int k = 0; k < 3; ++k) out[k] = c[k]; return; }
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) {
    double den = d2 - d6;
    double w = den != 0 ? d2 / den : 0.0;
    for (int k = 0; k < 3; ++k) out[k] = a[k] + w * ac[k];
    return;
  }
  double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    double den = (d4 - d3) + (d5 - d6);
    double w = den != 0 ? (d4 - d3) / den : 0.0;
    for (int k = 0; k < 3; ++k) out[k] = b[k] + w * (c[k] - b[k]);
    return;
  }
  double s = va + vb + vc;
  if (s == 0) {  // degenerate triangle: best of the three segments
    const double* E[3][2] = {{a, b}, {b, c}, {c, a}};
    double bestd = 1e300;
    for (int e = 0; e < 3; ++e) {
      double u[3], w[3];
      for (int k = 0; k < 3; ++k) { u[k] = E[e][1][k] - E[e][0][k]; w[k] = p[k] - E[e][0][k]; }
      double uu = dt(u, u), t = uu > 0 ? dt(u, w) / uu : 0.0;
      t = t < 0 ? 0 : (t > 1 ? 1 : t);
      double q[3], dd = 0;
      for (int k = 0; k < 3; ++k) { q[k] = E[e][0][k] + t * u[k]; dd += (p[k] - q[k]) * (p[k] - q[k]); }
      if (dd < bestd) { bestd = dd; for (int k = 0; k < 3; ++k) out[k] = q[k]; }
    }
    return;
  }
  double den = 1.0 / s, v = vb * den, w = vc * den;
  for (int k = 0; k < 3; ++k) out[k] = a[k] + ab[k] * v + ac[k] * w;
}

// Exact (FP64) unsigned distance by brute force over all faces.
void orc_fp64_udf(const float* V, const int32_t* F, int64_t nF, const float* P, int64_t K,
                  double* dist, double* closest, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t q = 0; q < K; q++) {
    double p[3] = {P[q * 3], P[q * 3 + 1], P[q * 3 + 2]};
    double best = 1e300, bc[3] = {p[0], p[1], p[2]};
    for (int64_t f = 0; f < nF; f++) {
      double a[3], b[3], c[3], o[3];
      for (int k = 0; k < 3; ++k) {
        a[k] = V[(int64_t)F[f * 3] * 3 + k]; b[k] = V[(int64_t)F[f * 3 + 1] * 3 + k];
        c[k] = V[(int64_t)F[f * 3 + 2] * 3 + k];
      }
      closest_fp64(a, b, c, p, o);
      double d2 = (o[0] - p[0]) * (o[0] - p[0]) + (o[1] - p[1]) * (o[1] - p[1]) + (o[2] - p[2]) * (o[2] - p[2]);
      if (d2 < best) { best = d2; bc[0] = o[0]; bc[1] = o[1]; bc[2] = o[2]; }
    }
    dist[q] = std::sqrt(best);
    closest[q * 3] = bc[0]; closest[q * 3 + 1] = bc[1]; closest[q * 3 + 2] = bc[2];
  }
}

// FP64 closest point of one (point, triangle) pair (for tests).
void orc_tri_closest_fp64(const double* tri9, const double* p, double* out) {
  closest_fp64(tri9, tri9 + 3, tri9 + 6, p, out);
}

// Reference AABB walk (bvh_kernels.cu:555-610) run sequentially: hit mask and
// the (box, face) pairs in the order one thread would append them.  Returns the
// total pair count; writes at most cap pairs.  canonical=1 additionally
// requires the triangle's own AABB to overlap the box (a tree-independent
// variant, used only to measure how far the reference's leaf-box rule
// reaches beyond the triangles' own boxes).
int64_t orc_bvh_aabb(const float* nodes_f, const float* tris, const float* qmin, const float* qmax,
                     int64_t K, uint8_t* hit, int32_t* pair_a, int32_t* pair_f, int64_t cap,
                     int canonical) {
  const Node* nodes = reinterpret_cast<const Node*>(nodes_f);
  int64_t n = 0;
  for (int64_t q = 0; q < K; q++) {
    f3 lo = ld3(qmin + q * 3), hi = ld3(qmax + q * 3);
    bool any = false;
    int stack[128]; int sp = 0; stack[sp++] = 0;
    while (sp > 0) {
      const Node& nd = nodes[stack[--sp]];
      if (!aabb_overlap(nd.bmin, nd.bmax, lo, hi)) continue;
      if (nd.left < 0) {
        int s = -nd.left - 1, e = -nd.right - 1;
        for (int i = s; i < e; i++) {
          const float* t = tris + (int64_t)i * 10;
          f3 a = ld3(t), b = ld3(t + 3), c = ld3(t + 6);
          if (canonical && !tri_box_overlap(a, b, c, lo, hi)) continue;
          if (tri_aabb_sat(a, b, c, lo, hi)) {
            any = true;
            if (n < cap) { pair_a[n] = (int32_t)q; std::memcpy(pair_f + n, t + 9, 4); }
            n++;
          }
        }
      } else if (sp < 127) {
        stack[sp++] = nd.right;
        stack[sp++] = nd.left;
      }
    }
    if (hit) hit[q] = any ? 1 : 0;
  }
  return n;
}

// Tree-independent hit mask: box hits iff some face has an overlapping AABB
// and passes the reference SAT.  Brute force over faces, parallel over boxes.
void orc_brute_aabb_hit(const float* V, const int32_t* F, int64_t nF, const float* qmin,
                        const float* qmax, int64_t K, uint8_t* hit, int32_t* count, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t q = 0; q < K; q++) {
    f3 lo = ld3(qmin + q * 3), hi = ld3(qmax + q * 3);
    int c = 0;
    for (int64_t f = 0; f < nF; f++) {
      f3 a = ld3(V + (int64_t)F[f * 3] * 3), b = ld3(V + (int64_t)F[f * 3 + 1] * 3),
         cc = ld3(V + (int64_t)F[f * 3 + 2] * 3);
      if (tri_box_overlap(a, b, cc, lo, hi) && tri_aabb_sat(a, b, cc, lo, hi)) c++;
    }
    hit[q] = c > 0;
    if (count) count[q] = c;
  }
}

// Walk-free statement of the reference's box query on a given tree: a
// (box, face) pair is a hit when the box overlaps the face's LEAF box and the
// face passes the SAT (bvh_kernels.cu:583-591).  Every leaf is scanned (no
// traversal, so a pruning mistake in a walk cannot hide here); node boxes
// contain their children's, so any leaf overlap implies the walk reaches it.
// Pass 1 (offsets == nullptr) writes counts[K]; pass 2 writes the pairs of box
// q in ascending face order at offsets[q].
void orc_leaf_brute_pairs(const float* nodes_f, int64_t n_nodes, const float* tris, const float* qmin,
                          const float* qmax, int64_t K, int32_t* counts, const int64_t* offsets, int32_t* pair_a,
                          int32_t* pair_f, int nthreads) {
  const Node* nodes = reinterpret_cast<const Node*>(nodes_f);
  std::vector<int64_t> leaves;
  for (int64_t i = 0; i < n_nodes; ++i)
    if (nodes[i].left < 0) leaves.push_back(i);
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t q = 0; q < K; q++) {
    f3 lo = ld3(qmin + q * 3), hi = ld3(qmax + q * 3);
    std::vector<int32_t> fs;
    for (int64_t li : leaves) {
      const Node& nd = nodes[li];
      if (!aabb_overlap(nd.bmin, nd.bmax, lo, hi)) continue;
      for (int i = -nd.left - 1; i < -nd.right - 1; i++) {
        const float* t = tris + (int64_t)i * 10;
        if (tri_aabb_sat(ld3(t), ld3(t + 3), ld3(t + 6), lo, hi)) {
          int32_t id;
          std::memcpy(&id, t + 9, 4);
          fs.push_back(id);
        }
      }
    }
    counts[q] = (int32_t)fs.size();
    if (offsets) {
      std::sort(fs.begin(), fs.end());
      for (size_t k = 0; k < fs.size(); ++k) {
        pair_a[offsets[q] + (int64_t)k] = (int32_t)q;
        pair_f[offsets[q] + (int64_t)k] = fs[k];
      }
    }
  }
}

// Cube boxes exactly as cube_aabb_level computes them in fp32
// (octree_indexer.py:331-336): min = float(ijk) * cell + b0, max = min + cell,
// cell = (b1 - b0) / 2^level.
static inline void cube_box(const int64_t* ijk, const float* b0, const float* cell, f3* lo, f3* hi) {
  float m[3], M[3];
  for (int k = 0; k < 3; ++k) {
    float t = (float)ijk[k] * cell[k];
    m[k] = t + b0[k];
    M[k] = m[k] + cell[k];
  }
  *lo = mk(m[0], m[1], m[2]);
  *hi = mk(M[0], M[1], M[2]);
}

// Octree level loop (octree_indexer.py:470-509) on a reference SAH tree.  A
// cube is kept when some leaf whose box overlaps the cube holds a triangle
// passing the SAT (bvh_kernels.cu:583-591); canonical = 1 additionally
// requires the triangle's own AABB to overlap (a tree-independent variant,
// kept only to measure the difference).  bounds = b0[3], b1[3].  Writes up
// to cap ijk rows (int64) at max_level in the reference order; returns the
// count.  tested_out (nullable) receives the number of cube tests done per
// level, hits_out the number of hits per level (max_level+1 entries each).
// full = 1 walks each cube's whole overlap set as the reference does (its
// octree calls intersect_aabb(mode=0), which enumerates every pair,
// bvh_kernels.cu:555-610) instead of stopping at the first hit; work_out
// (nullable, 2 entries) then receives the total node visits and triangle
// SAT tests of all cube tests.  Either way the kept set is the same.
int64_t orc_octree_traverse(const float* nodes_f, const float* tris, const float* bounds,
                            int min_level, int max_level, int64_t* out, int64_t cap,
                            int64_t* tested_out, int64_t* hits_out, int nthreads, int full,
                            int64_t* work_out, int canonical) {
  const Node* nodes = reinterpret_cast<const Node*>(nodes_f);
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const float* b0 = bounds;
  const float* b1 = bounds + 3;
  std::vector<int64_t> cur;
  int64_t res = (int64_t)1 << min_level;
  cur.reserve((size_t)(res * res * res * 3));
  for (int64_t i = 0; i < res; ++i)
    for (int64_t j = 0; j < res; ++j)
      for (int64_t k = 0; k < res; ++k) { cur.push_back(i); cur.push_back(j); cur.push_back(k); }
  static const int corners[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                    {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  for (int level = min_level; level <= max_level; ++level) {
    if (level > min_level) {
      if (cur.empty()) break;
      std::vector<int64_t> nxt(cur.size() * 8);
      for (size_t c = 0; c < cur.size() / 3; ++c)
        for (int o = 0; o < 8; ++o)
          for (int k = 0; k < 3; ++k) nxt[(c * 8 + o) * 3 + k] = cur[c * 3 + k] * 2 + corners[o][k];
      cur.swap(nxt);
    }
    float cell[3];
    float r = (float)((int64_t)1 << level);
    for (int k = 0; k < 3; ++k) cell[k] = (b1[k] - b0[k]) / r;
    int64_t n = (int64_t)cur.size() / 3;
    std::vector<uint8_t> hit((size_t)n);
    int64_t w_node = 0, w_tri = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : w_node, w_tri)
    for (int64_t q = 0; q < n; ++q) {
      f3 lo, hi;
      cube_box(&cur[(size_t)q * 3], b0, cell, &lo, &hi);
      bool any = false;
      int stack[128]; int sp = 0; stack[sp++] = 0;
      while (sp > 0 && (full || !any)) {
        const Node& nd = nodes[stack[--sp]];
        ++w_node;
        if (!aabb_overlap(nd.bmin, nd.bmax, lo, hi)) continue;
        if (nd.left < 0) {
          int s = -nd.left - 1, e = -nd.right - 1;
          for (int i = s; i < e && (full || !any); i++) {
            const float* t = tris + (int64_t)i * 10;
            f3 a = ld3(t), b = ld3(t + 3), c = ld3(t + 6);
            ++w_tri;
            if ((!canonical || tri_box_overlap(a, b, c, lo, hi)) && tri_aabb_sat(a, b, c, lo, hi)) any = true;
          }
        } else if (sp < 127) {
          stack[sp++] = nd.right;
          stack[sp++] = nd.left;
        }
      }
      hit[(size_t)q] = any;
    }
    std::vector<int64_t> kept;
    kept.reserve(cur.size());
    for (int64_t q = 0; q < n; ++q)
      if (hit[(size_t)q]) for (int k = 0; k < 3; ++k) kept.push_back(cur[(size_t)q * 3 + k]);
    if (tested_out) tested_out[level] = n;
    if (hits_out) hits_out[level] = (int64_t)kept.size() / 3;
    if (work_out) {
      work_out[0] += w_node;
      work_out[1] += w_tri;
    }
    cur.swap(kept);
  }
  int64_t n = (int64_t)cur.size() / 3;
  for (int64_t q = 0; q < n && q < cap; ++q)
    for (int k = 0; k < 3; ++k) out[q * 3 + k] = cur[(size_t)q * 3 + k];
  return n;
}

}  // extern "C"
