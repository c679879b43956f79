// Ray queries for gfx950: nearest hit, all hits, signed crossing counts.
//
// Replaces bvh_ray_intersect_kernel / bvh_ray_intersect_cuda and
// bvh_ray_all_hit_kernel / bvh_ray_all_hit_cuda (Atom3D
// atom3d/kernels/bvh_kernels.cu:469-550, 935-1022, 1068-1190) and the
// coalescing of MeshBVH._signed_crossing_count (atom3d/core/mesh_bvh.py:
// 906-970), on the blob built by bvh_build.hip.
//
// Per-triangle maths is the reference's: Moller-Trumbore with its FP64 u/v
// range tests (Triangle::ray_intersect, :185-202), the unit normal
// (Triangle::normal, :89-97) for the all-hit grazing filter and sign, and the
// barycentrics of Triangle::barycentric (:205-222).
//
// Results do not depend on the tree:
//   nearest hit = the minimum t over all faces (t < max_t, and not the
//   test's no-hit value 1e10), ties to the lowest face index;
//   all hits = every face passing the reference's filters, listed in (ray,
//   face) order after a3d_pairs_sort + a3d_ray_eval.
// Box tests widen every box by 2^-20 x (coordinate scale), which covers the
// FP32 rounding of the slab arithmetic, so a box holding a hit face is never
// skipped for rounding; the nearest-hit walk skips a box only when its entry
// distance exceeds the current best t by more than 1/128 of it.  (The
// FP32 t of a face can be off by more than that only for rays within ~1e-5
// rad of the face's plane; those faces could then be missed where the
// all-faces scan would take them.)
#include "a3d_common.hpp"

namespace a3d {
namespace {

constexpr int kBlk = 256;
// Per-lane LDS stack entries (the rest in private memory): 24 KB per block of the
// nearest-hit walk; its registers allow 5 wavefronts per SIMD, and 5 blocks' LDS fits a CU
// (8 entries: 6.75 ms, 12: 6.66 ms on 2^24 rays)
constexpr int kRayLds = 12;
constexpr int kRayOvf = kStackTotal - kRayLds;
constexpr float kNoHit = 1e10f;

// bvh_kernels.cu:185-202 (FP64 u/v tests as in the reference)
__device__ __forceinline__ float ray_tri(v3 a, v3 b, v3 c, v3 ro, v3 rd) {
  const float eps = 1e-8f;
  const double eps_edge = 1.19209290e-7;
  v3 e1 = sub3(b, a);
  v3 e2 = sub3(c, a);
  v3 h = cross3(rd, e2);
  float det = dot3(e1, h);
  if (fabsf(det) < eps) return kNoHit;
  float f = 1.0f / det;
  v3 s = sub3(ro, a);
  double u = (double)f * (double)dot3(s, h);
  if (u < -eps_edge || u > 1.0 + eps_edge) return kNoHit;
  v3 q = cross3(s, e1);
  double v = (double)f * (double)dot3(rd, q);
  if (v < -eps_edge || u + v > 1.0 + eps_edge) return kNoHit;
  float t = f * dot3(e2, q);
  return (t > eps) ? t : kNoHit;
}

// bvh_kernels.cu:89-97
__device__ __forceinline__ v3 tri_normal(v3 a, v3 b, v3 c) {
  v3 n = cross3(sub3(b, a), sub3(c, a));
  float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
  if (len > 1e-10f) return mul3(n, 1.0f / len);
  return mk(0.0f, 0.0f, 1.0f);
}

// All-hit filter and sign (bvh_kernels.cu:975-985): 0 when the face does not
// count, else sign(dot(normal, d)).
__device__ __forceinline__ int all_hit_sign(v3 a, v3 b, v3 c, v3 ro, v3 rd, float max_t, float& t) {
  t = ray_tri(a, b, c, ro, rd);
  if (!(t > 0.0f && t < max_t)) return 0;
  const float ndotd = dot3(tri_normal(a, b, c), rd);
  if (ndotd > -1e-4f && ndotd < 1e-4f) return 0;
  return ndotd > 0.0f ? 1 : -1;
}

// Reciprocal direction for the slab test; components below 1e-30 in
// magnitude are clamped (no 0 x inf).
__device__ __forceinline__ float safe_inv(float d) {
  return 1.0f / (fabsf(d) > 1e-30f ? d : (d >= 0.0f ? 1e-30f : -1e-30f));
}

// Entry distance (>= 0) of the ray into the box widened by w, or +inf
// (also for an empty, inverted box).
__device__ __forceinline__ float ray_box(v3 ro, v3 inv, float lx, float ly, float lz, float hx, float hy, float hz,
                                         float w) {
  if (!(lx <= hx)) return __int_as_float(0x7f800000);  // the empty child of a one-leaf tree
  float t1 = (lx - w - ro.x) * inv.x, t2 = (hx + w - ro.x) * inv.x;
  float tmin = fminf(t1, t2), tmax = fmaxf(t1, t2);
  t1 = (ly - w - ro.y) * inv.y; t2 = (hy + w - ro.y) * inv.y;
  tmin = fmaxf(tmin, fminf(t1, t2)); tmax = fminf(tmax, fmaxf(t1, t2));
  t1 = (lz - w - ro.z) * inv.z; t2 = (hz + w - ro.z) * inv.z;
  tmin = fmaxf(tmin, fminf(t1, t2)); tmax = fminf(tmax, fmaxf(t1, t2));
  return (tmax >= tmin && tmax >= 0.0f) ? fmaxf(tmin, 0.0f) : __int_as_float(0x7f800000);
}

__device__ __forceinline__ float box_widen(v3 ro, float coord_max) {
  const float m = fmaxf(fabsf(ro.x), fmaxf(fabsf(ro.y), fabsf(ro.z)));
  return 9.5367431640625e-07f * (coord_max + m);  // 2^-20
}

__device__ __forceinline__ float prune_t(float best) { return best + best * 0.0078125f; }  // x (1 + 2^-7)

// One step of a walk that keeps every hit (no pruning by t): visit the inner
// node cur (push the right child when both are entered), or report the leaf
// cur; pop = the walk continues from the stack.
__device__ __forceinline__ void all_step(const Node64* __restrict__ nodes, v3 ro, v3 inv, float wid, float lim,
                                         int32_t& cur, int& sp, int32_t* s_ref, int tid, int32_t* ovf_ref,
                                         bool& at_leaf, bool& pop) {
  if (cur >= 0) {
    const Node64* nd = nodes + cur;
    const float4 a = nd->a, b = nd->b, c = nd->c;
    const int4 d = nd->d;
    const bool vl = ray_box(ro, inv, a.x, a.y, a.z, a.w, b.x, b.y, wid) <= lim;
    const bool vr = ray_box(ro, inv, b.z, b.w, c.x, c.y, c.z, c.w, wid) <= lim;
    if (vl && vr) {
      if (sp < kRayLds) s_ref[sp * kBlk + tid] = d.y;
      else if (sp < kStackTotal) ovf_ref[sp - kRayLds] = d.y;
      sp = sp < kStackTotal ? sp + 1 : sp;
      cur = d.x;
    } else if (vl) {
      cur = d.x;
    } else if (vr) {
      cur = d.y;
    } else {
      pop = true;
    }
  } else {
    at_leaf = true;
    pop = true;
  }
}

// Pops the next node of a walk that keeps every hit; false when the stack is empty.
__device__ __forceinline__ bool all_pop(int32_t& cur, int& sp, const int32_t* s_ref, int tid, const int32_t* ovf_ref) {
  if (sp == 0) return false;
  --sp;
  cur = sp < kRayLds ? s_ref[sp * kBlk + tid] : ovf_ref[sp - kRayLds];
  return true;
}

// ------------------------------------------------------------ nearest hit
// One walk per ray.  The rays of a wavefront come from its LaneJobs list: a
// lane whose walk is over holds its result until enough lanes are free,
// then the free lanes write their results together and take the next rays
// (see lane_take).  Leaf triangles are dealt over the wavefront (see
// leaf_pair).  Each ray's walk takes the same steps and gives the same result
// as a walk on its own: a leaf's tests all complete before its lane pops,
// and the lowest (t, face id) wins as in a one-by-one scan.
__global__ __launch_bounds__(kBlk) void k_ray_nearest(const void* __restrict__ blob, const float* __restrict__ O,
                                                       const float* __restrict__ D, int64_t K, float max_t,
                                                       uint8_t* __restrict__ hit, float* __restrict__ t_out,
                                                       int32_t* __restrict__ face, float* __restrict__ pts,
                                                       float* __restrict__ uvw) {
  __shared__ int32_t s_ref[kRayLds * kBlk];
  __shared__ float s_t[kRayLds * kBlk];
  __shared__ unsigned long long s_key[kBlk];  // owner's best (t bits << 32 | face id) during a leaf round
  __shared__ int32_t s_tri[kBlk];             // triangle of the owner's winning pair
  __shared__ uint8_t s_own[kBlk];
  const int tid = threadIdx.x, lane = tid & 63;
  unsigned long long* key = s_key + (tid & ~63);
  int32_t* wtri = s_tri + (tid & ~63);
  uint8_t* own = s_own + (tid & ~63);
  const BvhHeader* H = header(blob);
  const Node64* __restrict__ nodes = nodes_of(blob);
  const Tri48* __restrict__ tris = tris_of(blob);
  const bool empty = H->n_tris == 0;
  LaneJobs J = lane_jobs(K, kBlk);

  int64_t r = -1;     // this lane's ray
  bool done = false;  // its walk is over, its result not yet written
  v3 ro = mk(0.0f, 0.0f, 0.0f), rd = ro, inv = ro;
  float wid = 0.0f, best = max_t, thr = 0.0f;
  int32_t best_id = 0x7fffffff, best_tri = -1;
  int32_t ovf_ref[kRayOvf];
  float ovf_t[kRayOvf];
  int sp = 0;
  int32_t cur = 0;
  while (true) {
    const uint64_t fm = __ballot(r < 0 || done);
    const int n_free = __popcll(fm);
    const bool more = J.next < J.njobs;
    if (n_free == 64 || (more && n_free >= kRefillIdle)) {
      if (done) {
        hit[r] = best_tri >= 0 ? 1 : 0;
        t_out[r] = best;
        face[r] = best_tri >= 0 ? best_id : -1;
        v3 p = ro, w = mk(0.0f, 0.0f, 0.0f);
        if (best_tri >= 0) {
          p = add3(ro, mul3(rd, best));
          v3 ta, tb, tc;
          int32_t id;
          load_tri(tris, best_tri, ta, tb, tc, id);
          w = tri_bary(ta, tb, tc, p);
        }
        if (pts) { pts[r * 3] = p.x; pts[r * 3 + 1] = p.y; pts[r * 3 + 2] = p.z; }
        if (uvw) { uvw[r * 3] = w.x; uvw[r * 3 + 1] = w.y; uvw[r * 3 + 2] = w.z; }
        done = false;
        r = -1;
      }
      if (!more) break;  // every lane free and no ray left
      const int64_t nr = lane_take(J, fm, K);
      if (nr >= 0) {
        r = nr;
        ro = mk(O[r * 3], O[r * 3 + 1], O[r * 3 + 2]);
        rd = mk(D[r * 3], D[r * 3 + 1], D[r * 3 + 2]);
        inv = mk(safe_inv(rd.x), safe_inv(rd.y), safe_inv(rd.z));
        wid = box_widen(ro, H->coord_max);
        best = max_t;
        best_id = 0x7fffffff;
        best_tri = -1;
        thr = prune_t(best);
        sp = 0;
        cur = 0;
      }
      continue;
    }
    bool pop = false, at_leaf = false;
    if (r >= 0 && !done) {
      if (empty) {
        pop = true;
      } else if (cur >= 0) {
        const Node64* nd = nodes + cur;
        const float4 a = nd->a, b = nd->b, c = nd->c;
        const int4 d = nd->d;
        const float tl = ray_box(ro, inv, a.x, a.y, a.z, a.w, b.x, b.y, wid);
        const float tr = ray_box(ro, inv, b.z, b.w, c.x, c.y, c.z, c.w, wid);
        const bool vl = tl <= thr, vr = tr <= thr;
        if (vl && vr) {
          int32_t near_ref = d.x, far_ref = d.y;
          float far_t = tr;
          if (tr < tl) { near_ref = d.y; far_ref = d.x; far_t = tl; }
          if (sp < kRayLds) {
            s_ref[sp * kBlk + tid] = far_ref;
            s_t[sp * kBlk + tid] = far_t;
          } else if (sp < kStackTotal) {
            ovf_ref[sp - kRayLds] = far_ref;
            ovf_t[sp - kRayLds] = far_t;
          }
          sp = sp < kStackTotal ? sp + 1 : sp;
          cur = near_ref;
        } else if (vl) {
          cur = d.x;
        } else if (vr) {
          cur = d.y;
        } else {
          pop = true;
        }
      } else {
        at_leaf = true;
        pop = true;
      }
    }
    if (__ballot(at_leaf)) {
      const LeafDeal dl = leaf_deal(at_leaf, cur, lane);
      for (int base = 0; base < dl.total; base += 64) {
        wave_fence();
        const unsigned long long mine = ((unsigned long long)__float_as_uint(best) << 32) | (uint32_t)best_id;
        key[lane] = mine;
        bool valid;
        int L;
        int32_t t;
        leaf_pair(dl, base, own, lane, valid, L, t);
        const v3 pro = mk(__shfl(ro.x, L), __shfl(ro.y, L), __shfl(ro.z, L));
        const v3 prd = mk(__shfl(rd.x, L), __shfl(rd.y, L), __shfl(rd.z, L));
        unsigned long long k = ~0ull;
        if (valid) {
          v3 ta, tb, tc;
          int32_t id;
          load_tri(tris, t, ta, tb, tc, id);
          const float th = ray_tri(ta, tb, tc, pro, prd);
          if (th != kNoHit && th < max_t) {
            k = ((unsigned long long)__float_as_uint(th) << 32) | (uint32_t)id;
            atomicMin(&key[L], k);
          }
        }
        wave_fence();
        if (valid && key[L] == k) wtri[L] = t;
        wave_fence();
        const unsigned long long nk = key[lane];
        if (nk != mine) {
          best = __uint_as_float((uint32_t)(nk >> 32));
          best_id = (int32_t)(uint32_t)nk;
          best_tri = wtri[lane];
          thr = prune_t(best);
        }
      }
    }
    if (!pop) continue;
    bool found = false;
    while (sp > 0) {
      --sp;
      int32_t rr;
      float tt;
      if (sp < kRayLds) {
        rr = s_ref[sp * kBlk + tid];
        tt = s_t[sp * kBlk + tid];
      } else {
        rr = ovf_ref[sp - kRayLds];
        tt = ovf_t[sp - kRayLds];
      }
      if (tt <= thr) { cur = rr; found = true; break; }
    }
    if (!found) done = true;
  }
}

// ------------------------------------------------------------ all hits
// offsets == nullptr: counts[r] = number of hits; else the hits of ray r are
// written from offsets[r] on, in walk order (leaves in the order the walk
// reaches them, a leaf's triangles in order): ray index, face, t, sign.
// Rays from the wavefront's LaneJobs list, leaf triangles dealt over the
// wavefront, as in k_ray_nearest; a hit's slot is its owner's position plus
// the owner's hits before it in the round.
__global__ __launch_bounds__(kBlk) void k_ray_all(const void* __restrict__ blob, const float* __restrict__ O,
                                                   const float* __restrict__ D, int64_t K, float max_t,
                                                   int32_t* __restrict__ counts, const int64_t* __restrict__ offsets,
                                                   int32_t* __restrict__ ray_out, int32_t* __restrict__ face_out,
                                                   float* __restrict__ t_out, int32_t* __restrict__ sign_out) {
  __shared__ int32_t s_ref[kRayLds * kBlk];
  __shared__ uint8_t s_own[kBlk];
  const int tid = threadIdx.x, lane = tid & 63;
  uint8_t* own = s_own + (tid & ~63);
  const BvhHeader* H = header(blob);
  const Node64* __restrict__ nodes = nodes_of(blob);
  const Tri48* __restrict__ tris = tris_of(blob);
  const bool empty = H->n_tris == 0;
  const float lim = prune_t(max_t);
  LaneJobs J = lane_jobs(K, kBlk);

  int64_t r = -1;
  bool done = false;
  v3 ro = mk(0.0f, 0.0f, 0.0f), rd = ro, inv = ro;
  float wid = 0.0f;
  int64_t pos = 0;
  int32_t n_hit = 0;
  int32_t ovf_ref[kRayOvf];
  int sp = 0;
  int32_t cur = 0;
  while (true) {
    const uint64_t fm = __ballot(r < 0 || done);
    const int n_free = __popcll(fm);
    const bool more = J.next < J.njobs;
    if (n_free == 64 || (more && n_free >= kRefillIdle)) {
      if (done) {
        if (!offsets) counts[r] = n_hit;
        done = false;
        r = -1;
      }
      if (!more) break;
      const int64_t nr = lane_take(J, fm, K);
      if (nr >= 0) {
        r = nr;
        ro = mk(O[r * 3], O[r * 3 + 1], O[r * 3 + 2]);
        rd = mk(D[r * 3], D[r * 3 + 1], D[r * 3 + 2]);
        inv = mk(safe_inv(rd.x), safe_inv(rd.y), safe_inv(rd.z));
        wid = box_widen(ro, H->coord_max);
        pos = offsets ? offsets[r] : 0;
        n_hit = 0;
        sp = 0;
        cur = 0;
      }
      continue;
    }
    bool pop = false, at_leaf = false;
    if (r >= 0 && !done) {
      if (empty) pop = true;
      else all_step(nodes, ro, inv, wid, lim, cur, sp, s_ref, tid, ovf_ref, at_leaf, pop);
    }
    if (__ballot(at_leaf)) {
      const LeafDeal dl = leaf_deal(at_leaf, cur, lane);
      for (int base = 0; base < dl.total; base += 64) {
        wave_fence();
        bool valid;
        int L;
        int32_t t;
        leaf_pair(dl, base, own, lane, valid, L, t);
        const v3 pro = mk(__shfl(ro.x, L), __shfl(ro.y, L), __shfl(ro.z, L));
        const v3 prd = mk(__shfl(rd.x, L), __shfl(rd.y, L), __shfl(rd.z, L));
        const long long oray = __shfl((long long)r, L);
        const long long opos = __shfl((long long)pos, L);
        const int first = __shfl(dl.excl, L) - base;  // the owner's first pair of the round, as a lane
        float th = 0.0f;
        int sg = 0;
        int32_t id = 0;
        if (valid) {
          v3 ta, tb, tc;
          load_tri(tris, t, ta, tb, tc, id);
          sg = all_hit_sign(ta, tb, tc, pro, prd, max_t, th);
        }
        const uint64_t hm = __ballot(sg != 0);
        if (offsets && sg) {
          const int64_t q = opos + __popcll(hm & lanes_below(lane) & ~lanes_below(first > 0 ? first : 0));
          ray_out[q] = (int32_t)oray;
          face_out[q] = id;
          if (t_out) t_out[q] = th;
          if (sign_out) sign_out[q] = sg;
        }
        const int c = __popcll(hm & own_pairs(dl, base));
        n_hit += c;
        pos += c;
      }
    }
    if (pop && !all_pop(cur, sp, s_ref, tid, ovf_ref)) done = true;
  }
}

// t, sign, point and barycentrics of (ray, face) pairs from the original mesh
// arrays (the face's vertices are the blob's, so t is the walk's t).
__global__ __launch_bounds__(kBlk) void k_ray_eval(const float* __restrict__ V, const int32_t* __restrict__ F,
                                                    const float* __restrict__ O, const float* __restrict__ D,
                                                    int64_t n, const int32_t* __restrict__ ray_idx,
                                                    const int32_t* __restrict__ face, float max_t,
                                                    float* __restrict__ t_out, int32_t* __restrict__ sign_out,
                                                    float* __restrict__ pts, float* __restrict__ uvw) {
  const int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i >= n) return;
  const int64_t r = ray_idx[i], f = face[i];
  const v3 ro = mk(O[r * 3], O[r * 3 + 1], O[r * 3 + 2]);
  const v3 rd = mk(D[r * 3], D[r * 3 + 1], D[r * 3 + 2]);
  const int32_t i0 = F[f * 3], i1 = F[f * 3 + 1], i2 = F[f * 3 + 2];
  const v3 a = mk(V[(int64_t)i0 * 3], V[(int64_t)i0 * 3 + 1], V[(int64_t)i0 * 3 + 2]);
  const v3 b = mk(V[(int64_t)i1 * 3], V[(int64_t)i1 * 3 + 1], V[(int64_t)i1 * 3 + 2]);
  const v3 c = mk(V[(int64_t)i2 * 3], V[(int64_t)i2 * 3 + 1], V[(int64_t)i2 * 3 + 2]);
  float t;
  const int sg = all_hit_sign(a, b, c, ro, rd, max_t, t);
  const v3 p = add3(ro, mul3(rd, t));
  const v3 w = tri_bary(a, b, c, p);
  if (t_out) t_out[i] = t;
  if (sign_out) sign_out[i] = sg;
  if (pts) { pts[i * 3] = p.x; pts[i * 3 + 1] = p.y; pts[i * 3 + 2] = p.z; }
  if (uvw) { uvw[i * 3] = w.x; uvw[i * 3 + 1] = w.y; uvw[i * 3 + 2] = w.z; }
}

// Signed crossing count per ray (mesh_bvh.py:906-970): the ray's hits are
// sorted by t in place (insertion sort of its segment), consecutive hits
// closer than coalesce_t form a group, an all-same-sign group counts as one
// crossing of that sign, a mixed group by the parity of its sign sum.
__global__ __launch_bounds__(kBlk) void k_ray_crossings(float* __restrict__ t, int32_t* __restrict__ sign,
                                                         const int64_t* __restrict__ offsets, int64_t K,
                                                         float coalesce_t, int32_t* __restrict__ out) {
  const int64_t r = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (r >= K) return;
  const int64_t b = offsets[r], e = offsets[r + 1];
  for (int64_t i = b + 1; i < e; ++i) {
    const float ti = t[i];
    const int32_t si = sign[i];
    int64_t j = i - 1;
    while (j >= b && (t[j] > ti || (t[j] == ti && sign[j] > si))) {
      t[j + 1] = t[j];
      sign[j + 1] = sign[j];
      --j;
    }
    t[j + 1] = ti;
    sign[j + 1] = si;
  }
  int32_t total = 0;
  int64_t i = b;
  while (i < e) {
    int32_t cnt = 1, sum = sign[i];
    int64_t j = i + 1;
    while (j < e && fabsf(t[j] - t[j - 1]) < coalesce_t) { ++cnt; sum += sign[j]; ++j; }
    const int32_t sg = (sum > 0) - (sum < 0);
    total += (abs(sum) == cnt) ? sg : (abs(sum) % 2) * sg;
    i = j;
  }
  out[r] = total;
}

// Signed crossing counts of N points along n_dirs directions in one walk per
// (point, direction): ray r = direction r / N from point r % N.  The ray's
// hits (the all-hit walk's, same filters) are kept in registers sorted by
// (t, sign) as they arrive, then grouped and counted exactly as
// k_ray_crossings does over a sorted list.  A ray with more than
// kCrossCap hits gets out[r] = kCrossOverflow and is left to the two-pass
// path (count, write, k_ray_crossings).
constexpr int kCrossCap = 8;
constexpr int32_t kCrossOverflow = INT32_MIN;

__device__ __forceinline__ int32_t cross_total(const float (&hk)[kCrossCap], int32_t n_hit, float coalesce_t) {
  if (n_hit > kCrossCap) return kCrossOverflow;
  // groups of consecutive hits closer than coalesce_t (k_ray_crossings' rule)
  int32_t total = 0, cnt = 0, sum = 0;
  float prev = 0.0f;
#pragma unroll
  for (int k = 0; k < kCrossCap; ++k) {
    if (k < n_hit) {
      const float tk = fabsf(hk[k]);
      if (cnt > 0 && !(fabsf(tk - prev) < coalesce_t)) {
        const int32_t g = (sum > 0) - (sum < 0);
        total += (abs(sum) == cnt) ? g : (abs(sum) % 2) * g;
        cnt = 0;
        sum = 0;
      }
      ++cnt;
      sum += hk[k] > 0.0f ? 1 : -1;
      prev = tk;
    }
  }
  if (cnt > 0) {
    const int32_t g = (sum > 0) - (sum < 0);
    total += (abs(sum) == cnt) ? g : (abs(sum) % 2) * g;
  }
  return total;
}

// Rays from the wavefront's LaneJobs list, leaf triangles dealt over the
// wavefront, as in k_ray_nearest; each owner inserts its pairs' hits.
__global__ __launch_bounds__(kBlk) void k_ray_cross(const void* __restrict__ blob, const float* __restrict__ P,
                                                     int64_t N, const float* __restrict__ dirs, int64_t K,
                                                     float max_t, float coalesce_t, int32_t* __restrict__ out) {
  __shared__ int32_t s_ref[kRayLds * kBlk];
  __shared__ uint8_t s_own[kBlk];
  __shared__ float s_hv[kBlk];  // the signed t of each pair of a round (0: no hit)
  const int tid = threadIdx.x, lane = tid & 63;
  uint8_t* own = s_own + (tid & ~63);
  float* hv = s_hv + (tid & ~63);
  const BvhHeader* H = header(blob);
  const Node64* __restrict__ nodes = nodes_of(blob);
  const Tri48* __restrict__ tris = tris_of(blob);
  const bool empty = H->n_tris == 0;
  const float lim = prune_t(max_t);
  LaneJobs J = lane_jobs(K, kBlk);

  int64_t r = -1;
  bool done = false;
  v3 ro = mk(0.0f, 0.0f, 0.0f), rd = ro, inv = ro;
  float wid = 0.0f;
  // a hit is kept as sign * t (t > 0): |h| is its t exactly, and at equal t
  // the negative sign orders first, as in k_ray_crossings' (t, sign) order
  float hk[kCrossCap];
#pragma unroll
  for (int k = 0; k < kCrossCap; ++k) hk[k] = __int_as_float(0x7f800000);
  int32_t n_hit = 0;
  int32_t ovf_ref[kRayOvf];
  int sp = 0;
  int32_t cur = 0;
  while (true) {
    const uint64_t fm = __ballot(r < 0 || done);
    const int n_free = __popcll(fm);
    const bool more = J.next < J.njobs;
    if (n_free == 64 || (more && n_free >= kRefillIdle)) {
      if (done) {
        out[r] = cross_total(hk, n_hit, coalesce_t);
        done = false;
        r = -1;
      }
      if (!more) break;
      const int64_t nr = lane_take(J, fm, K);
      if (nr >= 0) {
        r = nr;
        const int64_t pi = r % N, di = r / N;
        ro = mk(P[pi * 3], P[pi * 3 + 1], P[pi * 3 + 2]);
        rd = mk(dirs[di * 3], dirs[di * 3 + 1], dirs[di * 3 + 2]);
        inv = mk(safe_inv(rd.x), safe_inv(rd.y), safe_inv(rd.z));
        wid = box_widen(ro, H->coord_max);
#pragma unroll
        for (int k = 0; k < kCrossCap; ++k) hk[k] = __int_as_float(0x7f800000);
        n_hit = 0;
        sp = 0;
        cur = 0;
      }
      continue;
    }
    bool pop = false, at_leaf = false;
    if (r >= 0 && !done) {
      if (empty) pop = true;
      else all_step(nodes, ro, inv, wid, lim, cur, sp, s_ref, tid, ovf_ref, at_leaf, pop);
    }
    if (__ballot(at_leaf)) {
      const LeafDeal dl = leaf_deal(at_leaf, cur, lane);
      for (int base = 0; base < dl.total; base += 64) {
        wave_fence();
        bool valid;
        int L;
        int32_t t;
        leaf_pair(dl, base, own, lane, valid, L, t);
        const v3 pro = mk(__shfl(ro.x, L), __shfl(ro.y, L), __shfl(ro.z, L));
        const v3 prd = mk(__shfl(rd.x, L), __shfl(rd.y, L), __shfl(rd.z, L));
        float x = 0.0f;
        int sg = 0;
        if (valid) {
          v3 ta, tb, tc;
          int32_t id;
          load_tri(tris, t, ta, tb, tc, id);
          float th;
          sg = all_hit_sign(ta, tb, tc, pro, prd, max_t, th);
          x = sg > 0 ? th : -th;
        }
        hv[lane] = x;
        const uint64_t hm = __ballot(sg != 0);
        wave_fence();
        uint64_t mine = hm & own_pairs(dl, base);
        n_hit += __popcll(mine);
        while (mine) {
          const int k = __builtin_ctzll(mine);
          mine &= mine - 1ull;
          // sorted insertion: the new hit moves down past every kept hit it precedes
          float y = hv[k];
#pragma unroll
          for (int m = 0; m < kCrossCap; ++m) {
            const float ay = fabsf(y), ah = fabsf(hk[m]);
            const bool before = ay < ah || (ay == ah && y < hk[m]);
            const float o = hk[m];
            hk[m] = before ? y : o;
            y = before ? o : y;
          }
        }
      }
    }
    if (pop && !all_pop(cur, sp, s_ref, tid, ovf_ref)) done = true;
  }
}

inline unsigned grid_for(int64_t n) { return (unsigned)div_up(n > 0 ? n : 1, kBlk); }

int check_rays(const char* fn, const void* blob, const float* O, const float* D, int64_t K) {
  if (K < 0 || K > (int64_t(1) << 31) - 1) return fail(A3D_ERR_ARG, "%s: K=%lld out of range [0, 2^31)", fn, (long long)K);
  if (K > 0 && (!blob || !O || !D)) return fail(A3D_ERR_ARG, "%s: null blob or ray arrays", fn);
  return A3D_OK;
}

}  // namespace
}  // namespace a3d

using namespace a3d;

extern "C" int a3d_ray_nearest(const void* blob, const float* O, const float* D, int64_t K, float max_t,
                               uint8_t* hit, float* t, int32_t* face, float* points, float* uvw,
                               a3d_stream_t stream) {
  int rc = check_rays("a3d_ray_nearest", blob, O, D, K);
  if (rc) return rc;
  if (K == 0) return A3D_OK;
  if (!hit || !t || !face) return fail(A3D_ERR_ARG, "a3d_ray_nearest: null output");
  static int resident[64];
  hipLaunchKernelGGL(k_ray_nearest, dim3(refill_grid(k_ray_nearest, kBlk, K, resident)), dim3(kBlk), 0,
                     (hipStream_t)stream, blob, O, D, K, max_t, hit, t, face, points, uvw);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_ray_all_count(const void* blob, const float* O, const float* D, int64_t K, float max_t,
                                 int32_t* counts, a3d_stream_t stream) {
  int rc = check_rays("a3d_ray_all_count", blob, O, D, K);
  if (rc) return rc;
  if (K == 0) return A3D_OK;
  if (!counts) return fail(A3D_ERR_ARG, "a3d_ray_all_count: null counts");
  static int resident[64];
  hipLaunchKernelGGL(k_ray_all, dim3(refill_grid(k_ray_all, kBlk, K, resident)), dim3(kBlk), 0, (hipStream_t)stream,
                     blob, O, D, K, max_t, counts, (const int64_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr,
                     (float*)nullptr, (int32_t*)nullptr);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_ray_all_hits(const void* blob, const float* O, const float* D, int64_t K, float max_t,
                                const int64_t* offsets, int32_t* ray_idx, int32_t* face, float* t, int32_t* sign,
                                a3d_stream_t stream) {
  int rc = check_rays("a3d_ray_all_hits", blob, O, D, K);
  if (rc) return rc;
  if (K == 0) return A3D_OK;
  if (!offsets || !ray_idx || !face) return fail(A3D_ERR_ARG, "a3d_ray_all_hits: null offsets or outputs");
  static int resident[64];
  hipLaunchKernelGGL(k_ray_all, dim3(refill_grid(k_ray_all, kBlk, K, resident)), dim3(kBlk), 0, (hipStream_t)stream,
                     blob, O, D, K, max_t, (int32_t*)nullptr, offsets, ray_idx, face, t, sign);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_ray_eval(const float* V, int64_t nV, const int32_t* F, int64_t nF, const float* O,
                            const float* D, int64_t n, const int32_t* ray_idx, const int32_t* face, float max_t,
                            float* t, int32_t* sign, float* points, float* uvw, a3d_stream_t stream) {
  if (n < 0) return fail(A3D_ERR_ARG, "a3d_ray_eval: n=%lld < 0", (long long)n);
  if (n == 0) return A3D_OK;
  if (!V || !F || !O || !D || !ray_idx || !face || nV <= 0 || nF <= 0)
    return fail(A3D_ERR_ARG, "a3d_ray_eval: null or empty input");
  hipLaunchKernelGGL(k_ray_eval, dim3(grid_for(n)), dim3(kBlk), 0, (hipStream_t)stream, V, F, O, D, n, ray_idx, face,
                     max_t, t, sign, points, uvw);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_ray_crossings_dirs(const void* blob, const float* P, int64_t N, const float* dirs, int n_dirs,
                                      float max_t, float coalesce_t, int32_t* out, a3d_stream_t stream) {
  if (N < 0 || n_dirs < 0 || n_dirs > 4096 || N * n_dirs > (int64_t(1) << 31) - 1)
    return fail(A3D_ERR_ARG, "a3d_ray_crossings_dirs: N=%lld x n_dirs=%d out of range", (long long)N, n_dirs);
  const int64_t K = N * n_dirs;
  if (K == 0) return A3D_OK;
  if (!blob || !P || !dirs || !out) return fail(A3D_ERR_ARG, "a3d_ray_crossings_dirs: null argument");
  static int resident[64];
  hipLaunchKernelGGL(k_ray_cross, dim3(refill_grid(k_ray_cross, kBlk, K, resident)), dim3(kBlk), 0,
                     (hipStream_t)stream, blob, P, N, dirs, K, max_t, coalesce_t, out);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_ray_crossings(float* t, int32_t* sign, const int64_t* offsets, int64_t K, float coalesce_t,
                                 int32_t* out, a3d_stream_t stream) {
  if (K < 0) return fail(A3D_ERR_ARG, "a3d_ray_crossings: K=%lld < 0", (long long)K);
  if (K == 0) return A3D_OK;
  if (!offsets || !out) return fail(A3D_ERR_ARG, "a3d_ray_crossings: null offsets or output");
  hipLaunchKernelGGL(k_ray_crossings, dim3(grid_for(K)), dim3(kBlk), 0, (hipStream_t)stream, t, sign, offsets, K,
                     coalesce_t, out);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}
