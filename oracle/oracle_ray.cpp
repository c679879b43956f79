// Oracle — CPU restatement of the reference's ray queries.
//
// TEST INFRASTRUCTURE ONLY: the checker for the ray entry points of the HIP
// library (atom3d_amd/csrc/ray.hip).  Linked by tests/ and never by the
// product path.  Plain C++, -ffp-contract=off, IEEE division and sqrt, so
// every FP32 (and the reference's FP64) operation rounds as the reference
// writes it.
//
// Restated from (Atom3D, atom3d/kernels/bvh_kernels.cu):
//   Triangle::normal                 :89-97
//   Triangle::ray_intersect          :185-202  (Moller-Trumbore, FP64 u/v tests)
//   Triangle::barycentric            :205-222
//   ray_aabb_intersect               :256-281
//   bvh_ray_intersect_kernel         :469-550  (nearest hit, SAH-tree walk)
//   bvh_ray_all_hit_kernel           :935-1022 (all hits, grazing filter, sign)
// and atom3d/core/mesh_bvh.py _signed_crossing_count :906-970 (coalescing of
// near-coincident hits, per-group contribution).
//
// Tree-independent definitions used as the parity target (the reference's
// answers depend on its tree only through ties and pruning):
//   nearest hit = minimum t over ALL faces of ray_intersect, ties to the
//   lowest face index; a face counts when t < max_t and t is not the no-hit
//   value 1e10 (the reference also takes 1e10 as a hit when max_t > 1e10);
//   all hits = every face passing the kernel's filters, listed per ray in
//   face order.
//
// Pinning (tests/test_oracle_ray.py): the ray-triangle t and hit decisions
// against an FP64 plane intersection + inside test on random rays (away from
// edges) and against hand-made cases (axis rays through a unit triangle,
// parallel rays, rays behind the origin, edge and vertex hits within
// FLT_EPSILON); nearest hits of the all-faces scan against the reference's
// own SAH walk; crossing counts against closed forms (points inside / outside
// closed spheres give odd / even counts along every direction).
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
static inline f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

constexpr float kNoHit = 1e10f;

// bvh_kernels.cu:185-202
static float ray_tri(f3 a, f3 b, f3 c, f3 ro, f3 rd) {
  const float eps = 1e-8f;
  const double eps_edge = 1.19209290e-7;
  f3 e1 = sub3(b, a);
  f3 e2 = sub3(c, a);
  f3 h = cross3(rd, e2);
  float det = dot3(e1, h);
  if (fabsf(det) < eps) return kNoHit;
  float f = 1.0f / det;
  f3 s = sub3(ro, a);
  double u = (double)f * (double)dot3(s, h);
  if (u < -eps_edge || u > 1.0 + eps_edge) return kNoHit;
  f3 q = cross3(s, e1);
  double v = (double)f * (double)dot3(rd, q);
  if (v < -eps_edge || u + v > 1.0 + eps_edge) return kNoHit;
  float t = f * dot3(e2, q);
  return (t > eps) ? t : kNoHit;
}

// bvh_kernels.cu:89-97
static f3 tri_normal(f3 a, f3 b, f3 c) {
  f3 n = cross3(sub3(b, a), sub3(c, a));
  float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
  if (len > 1e-10f) return mul3(n, 1.0f / len);
  return mk(0.0f, 0.0f, 1.0f);
}

// bvh_kernels.cu:205-222
static f3 tri_bary(f3 a, f3 b, f3 c, f3 p) {
  f3 v0 = sub3(b, a), v1 = sub3(c, a), v2 = sub3(p, a);
  float d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1);
  float d20 = dot3(v2, v0), d21 = dot3(v2, v1);
  float denom = d00 * d11 - d01 * d01;
  float v = (d11 * d20 - d01 * d21) / fmaxf(denom, 1e-10f);
  float w = (d00 * d21 - d01 * d20) / fmaxf(denom, 1e-10f);
  return mk(1.0f - v - w, v, w);
}

// bvh_kernels.cu:256-281
static float ray_box(f3 ro, f3 rd, const float* bmin, const float* bmax) {
  f3 inv = mk(1.0f / (fabsf(rd.x) > 1e-8f ? rd.x : (rd.x >= 0 ? 1e-8f : -1e-8f)),
              1.0f / (fabsf(rd.y) > 1e-8f ? rd.y : (rd.y >= 0 ? 1e-8f : -1e-8f)),
              1.0f / (fabsf(rd.z) > 1e-8f ? rd.z : (rd.z >= 0 ? 1e-8f : -1e-8f)));
  float t1 = (bmin[0] - ro.x) * inv.x, t2 = (bmax[0] - ro.x) * inv.x;
  float tmin = fminf(t1, t2), tmax = fmaxf(t1, t2);
  t1 = (bmin[1] - ro.y) * inv.y; t2 = (bmax[1] - ro.y) * inv.y;
  tmin = fmaxf(tmin, fminf(t1, t2)); tmax = fminf(tmax, fmaxf(t1, t2));
  t1 = (bmin[2] - ro.z) * inv.z; t2 = (bmax[2] - ro.z) * inv.z;
  tmin = fmaxf(tmin, fminf(t1, t2)); tmax = fminf(tmax, fmaxf(t1, t2));
  if (tmax >= tmin && tmax >= 0) return fmaxf(tmin, 0.0f);
  return kNoHit;
}

struct Node { float bmin[3]; float bmax[3]; int left, right, pad; };

static inline void face_verts(const float* V, const int32_t* F, int64_t f, f3& a, f3& b, f3& c) {
  a = ld3(V + (int64_t)F[f * 3] * 3);
  b = ld3(V + (int64_t)F[f * 3 + 1] * 3);
  c = ld3(V + (int64_t)F[f * 3 + 2] * 3);
}

// All-hit filter of bvh_kernels.cu:975-985: 0 for no hit, else the sign.
static inline int all_hit_sign(f3 a, f3 b, f3 c, f3 ro, f3 rd, float max_t, float& t) {
  t = ray_tri(a, b, c, ro, rd);
  if (!(t > 0.0f && t < max_t)) return 0;
  float ndotd = dot3(tri_normal(a, b, c), rd);
  const float graze_eps = 1e-4f;
  if (ndotd > -graze_eps && ndotd < graze_eps) return 0;
  return ndotd > 0.0f ? 1 : -1;
}

}  // namespace

extern "C" {

// Ray-triangle t for one (ray, triangle) pair (bvh_kernels.cu:185-202);
// 1e10 when there is no hit.  Also the normal (:89-97).
float orc_ray_tri(const float* tri9, const float* ro, const float* rd, float* normal_out) {
  f3 a = ld3(tri9), b = ld3(tri9 + 3), c = ld3(tri9 + 6);
  if (normal_out) {
    f3 n = tri_normal(a, b, c);
    normal_out[0] = n.x; normal_out[1] = n.y; normal_out[2] = n.z;
  }
  return ray_tri(a, b, c, ld3(ro), ld3(rd));
}

// Nearest hit by all-faces scan: t < max_t, not the no-hit value, minimum t,
// ties to the lowest face.  hit_t = max_t and face = -1 when nothing is hit;
// points = o + d t and uvw = barycentric(point) (zeros) when hit (missed).
void orc_brute_ray(const float* V, const int32_t* F, int64_t nF, const float* O, const float* D, int64_t K,
                   float max_t, uint8_t* hit, float* hit_t, int32_t* face, float* pts, float* uvw, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t r = 0; r < K; ++r) {
    f3 ro = ld3(O + r * 3), rd = ld3(D + r * 3);
    float best = max_t;
    int64_t bf = -1;
    for (int64_t f = 0; f < nF; ++f) {
      f3 a, b, c;
      face_verts(V, F, f, a, b, c);
      float t = ray_tri(a, b, c, ro, rd);
      if (t != kNoHit && t < best) { best = t; bf = f; }
    }
    hit[r] = bf >= 0;
    hit_t[r] = best;
    face[r] = (int32_t)bf;
    f3 p = ro, w = mk(0.0f, 0.0f, 0.0f);
    if (bf >= 0) {
      p = add3(ro, mul3(rd, best));
      f3 a, b, c;
      face_verts(V, F, bf, a, b, c);
      w = tri_bary(a, b, c, p);
    }
    pts[r * 3] = p.x; pts[r * 3 + 1] = p.y; pts[r * 3 + 2] = p.z;
    uvw[r * 3] = w.x; uvw[r * 3 + 1] = w.y; uvw[r * 3 + 2] = w.z;
  }
}

// The reference's own nearest-hit walk (bvh_kernels.cu:469-550) over a tree
// from orc_sah_build (nodes [Nn, 9], tris [F, 10]).  work (nullable, int64[2]):
// node fetches and triangle tests summed over the rays.
void orc_sah_ray(const float* nodes_f, const float* tris, const float* O, const float* D, int64_t K, float max_t,
                 uint8_t* hit, float* hit_t, int32_t* face, int threads, int64_t* work) {
  int64_t n_node = 0, n_tri = 0;
  const Node* nodes = reinterpret_cast<const Node*>(nodes_f);
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : n_node, n_tri)
  for (int64_t r = 0; r < K; ++r) {
    f3 ro = ld3(O + r * 3), rd = ld3(D + r * 3);
    float mint = max_t;
    int best_face = -1;
    int stack[96];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const Node& nd = nodes[stack[--sp]];
      ++n_node;
      if (ray_box(ro, rd, nd.bmin, nd.bmax) >= mint) continue;
      if (nd.left < 0) {
        for (int i = -nd.left - 1; i < -nd.right - 1; ++i) {
          ++n_tri;
          const float* t9 = tris + (int64_t)i * 10;
          float t = ray_tri(ld3(t9), ld3(t9 + 3), ld3(t9 + 6), ro, rd);
          if (t < mint) {
            mint = t;
            std::memcpy(&best_face, t9 + 9, 4);
          }
        }
      } else {
        float t1 = ray_box(ro, rd, nodes[nd.left].bmin, nodes[nd.left].bmax);
        float t2 = ray_box(ro, rd, nodes[nd.right].bmin, nodes[nd.right].bmax);
        if (t1 < t2) {
          if (t2 < mint && sp < 96) stack[sp++] = nd.right;
          if (t1 < mint && sp < 96) stack[sp++] = nd.left;
        } else {
          if (t1 < mint && sp < 96) stack[sp++] = nd.left;
          if (t2 < mint && sp < 96) stack[sp++] = nd.right;
        }
      }
    }
    hit[r] = best_face >= 0;
    hit_t[r] = mint;
    face[r] = best_face;
  }
  if (work) { work[0] = n_node; work[1] = n_tri; }
}

// All hits of every ray, listed per ray in face order.  offsets == nullptr:
// writes counts[K] only; else writes, from offsets[r] on, the hits' ray
// index, face, t, sign, point (o + d t) and barycentrics.
void orc_brute_ray_all(const float* V, const int32_t* F, int64_t nF, const float* O, const float* D, int64_t K,
                       float max_t, int64_t* counts, const int64_t* offsets, int32_t* ray_idx, int32_t* face,
                       float* hit_t, int32_t* sign, float* pts, float* uvw, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t r = 0; r < K; ++r) {
    f3 ro = ld3(O + r * 3), rd = ld3(D + r * 3);
    int64_t n = 0;
    for (int64_t f = 0; f < nF; ++f) {
      f3 a, b, c;
      face_verts(V, F, f, a, b, c);
      float t;
      int s = all_hit_sign(a, b, c, ro, rd, max_t, t);
      if (!s) continue;
      if (offsets) {
        int64_t o = offsets[r] + n;
        f3 p = add3(ro, mul3(rd, t));
        f3 w = tri_bary(a, b, c, p);
        ray_idx[o] = (int32_t)r; face[o] = (int32_t)f; hit_t[o] = t; sign[o] = s;
        pts[o * 3] = p.x; pts[o * 3 + 1] = p.y; pts[o * 3 + 2] = p.z;
        uvw[o * 3] = w.x; uvw[o * 3 + 1] = w.y; uvw[o * 3 + 2] = w.z;
      }
      ++n;
    }
    if (!offsets) counts[r] = n;
  }
}

// Signed crossing count per ray (mesh_bvh.py:906-970) from hit lists
// (ray-major, offsets[K+1]): hits sorted by t, consecutive hits closer than
// coalesce_t merged into one group; a group whose signs all agree counts as
// one crossing of that sign, a mixed group by the parity of its sign sum.
void orc_crossings(const float* hit_t, const int32_t* sign, const int64_t* offsets, int64_t K, float coalesce_t,
                   int32_t* out) {
  for (int64_t r = 0; r < K; ++r) {
    std::vector<std::pair<float, int>> h;
    for (int64_t i = offsets[r]; i < offsets[r + 1]; ++i) h.emplace_back(hit_t[i], sign[i]);
    std::sort(h.begin(), h.end());
    int32_t total = 0;
    size_t i = 0;
    while (i < h.size()) {
      int cnt = 1, sum = h[i].second;
      size_t j = i + 1;
      while (j < h.size() && fabsf(h[j].first - h[j - 1].first) < coalesce_t) { ++cnt; sum += h[j].second; ++j; }
      int sg = (sum > 0) - (sum < 0);
      total += (std::abs(sum) == cnt) ? sg : (std::abs(sum) % 2) * sg;
      i = j;
    }
    out[r] = total;
  }
}

}  // extern "C"
