// Closest-point (UDF) queries for gfx950.
//
// Replaces bvh_udf_kernel + bvh_udf_cuda (Atom3D
// atom3d/kernels/bvh_kernels.cu:363-463, 1028-1066), which walk the tree with
// one thread per query, a 128-entry private stack and 36-byte AoS nodes.
//
// Default path (wave-coherent):
//   1. queries are ordered along a 30-bit Morton curve of their position
//      (radix sort of (key, index)), so the 64 queries of a wavefront are
//      spatial neighbours;
//   2. each wavefront walks the tree as one packet: a node or triangle is
//      fetched once per wave with scalar loads (64 bytes of node, 48 bytes of
//      triangle) and every lane tests it against its own query; a child is
//      entered when any lane's pruning test accepts it, nearer-by-vote first;
//      the stack (reference + child box, 32 bytes an entry) is shared by the
//      wave in LDS and re-tested against every lane when popped;
//   3. each lane writes its results to the query's original position.
// Fallback path (flag A3D_UDF_LANE): one independent walk per lane with a
// per-lane LDS stack.
//
// Closest-face definition (tree and order independent): the face minimising
// the reference's FP32 squared distance (Triangle::distance_sq, :101-135),
// ties to the lowest original face id.  A box is skipped only when its
// distance exceeds the lane's best distance by more than the query's FP32
// error slack (and thin-triangle error is folded into the leaf boxes at build
// time), so no face the formula ranks first can be skipped; the minimum over
// the tested faces is therefore the minimum over all faces, whatever order or
// extra nodes a wave visits.
#include "a3d_prim.hpp"

#include <mutex>

#include "a3d_common.hpp"

namespace a3d {
namespace {

constexpr int kBlk = 256;
constexpr int kWaves = kBlk / 64;
// The packet walk runs one wavefront per block: a block's LDS (stack and
// queue) is then released as soon as its own walk ends instead of waiting
// for the slowest of several wavefronts (occupancy 64 -> 74 %, the 4.1M-face
// walk 2.5 % faster).
constexpr int kBlkW = 64;
constexpr int kLaneLds = 16;  // per-lane LDS stack entries in the lane path
constexpr int kLaneOvf = kStackTotal - kLaneLds;

__device__ __forceinline__ int32_t uni(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ void write_result(int64_t q, v3 p, float best, int32_t best_id, int32_t best_tri,
                                             const Tri48* __restrict__ tris, float* __restrict__ dist,
                                             int32_t* __restrict__ face, float* __restrict__ closest,
                                             float* __restrict__ uvw) {
  dist[q] = sqrtf(best);
  face[q] = best_tri >= 0 ? best_id : -1;
  v3 cp = p, w = mk(0.0f, 0.0f, 0.0f);
  if (best_tri >= 0) {
    v3 ta, tb, tc;
    int32_t id;
    load_tri(tris, best_tri, ta, tb, tc, id);
    cp = tri_closest(ta, tb, tc, p);
    w = tri_bary(ta, tb, tc, cp);
  }
  closest[q * 3] = cp.x;
  closest[q * 3 + 1] = cp.y;
  closest[q * 3 + 2] = cp.z;
  if (uvw) {
    uvw[q * 3] = w.x;
    uvw[q * 3 + 1] = w.y;
    uvw[q * 3 + 2] = w.z;
  }
}

// ------------------------------------------------------------ exact-test queue
// The exact point-triangle tests of a wave are compacted: at a leaf, every
// (lane, triangle) pair whose box bound passes the lane's threshold is
// appended to a per-wave LDS queue.  Pairs are first filtered 64 at a time,
// one pair per lane, by the triangle's oriented box against the owning
// lane's current threshold; the survivors stay at the front of the queue and
// each batch of 64 of them is tested exactly, again one pair per lane.  A
// lane's best (squared distance, face id) is the 64-bit minimum of its old
// best and its tested pairs (float bits of a non-negative distance order like
// the distance, ties to the lowest id: the same rule as testing them one by
// one).  Deferring the tests only delays the tightening of the pruning
// threshold, which stays conservative.
//
// Queue layout: [0, qf) filtered, [qf, qn) not yet filtered.  Between leaf
// triangles qf < 64 and qn - qf < 64, and one triangle adds at most 64
// pairs, so qn < 192.
constexpr int kQCap = 192;
constexpr int kWaveStack = 100;  // >= 97 (see kStackTotal); sized so 8 blocks of 4 waves fit in 160 KiB of LDS
static_assert(kWaveStack >= kMaxWalkDepth + 1 && kStackTotal >= kMaxWalkDepth + 1,
              "walk stacks must hold every tree a3d_bvh_build accepts");

struct WaveQ {
  float4 ob[6];                // the current node's Obb96, copied by the node visit (global -> LDS)
  unsigned long long key[64];  // per-lane best key during a flush
  int32_t tri[64];             // triangle index of the winning pair
  int32_t qt[kQCap];           // queued pairs: triangle index
  uint8_t ql[kQCap];           //               and lane
};

// s_waitcnt immediate (gfx9 encoding) waiting for every vector memory access: vmcnt 0
// (bits 3:0 and 15:14), expcnt 7 and lgkmcnt 15 (no wait)
constexpr int kWaitVm0 = 0x0F70;

__device__ __forceinline__ unsigned long long best_key(float d2, int32_t id) {
  return ((unsigned long long)__float_as_uint(d2) << 32) | (uint32_t)id;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Moves queue entries [from, from + cnt) to [to, to + cnt) (to <= from,
// cnt <= 128).  Ends with the wave synchronised.
__device__ __forceinline__ void q_move(WaveQ* __restrict__ Q, int from, int to, int cnt, int lane) {
  const bool a = lane < cnt, b = lane + 64 < cnt;
  const int32_t ta = a ? Q->qt[from + lane] : 0, tb = b ? Q->qt[from + 64 + lane] : 0;
  const uint8_t la = a ? Q->ql[from + lane] : (uint8_t)0, lb = b ? Q->ql[from + 64 + lane] : (uint8_t)0;
  wave_sync();
  if (a) {
    Q->qt[to + lane] = ta;
    Q->ql[to + lane] = la;
  }
  if (b) {
    Q->qt[to + 64 + lane] = tb;
    Q->ql[to + 64 + lane] = lb;
  }
  wave_sync();
}

// Filters the n (1..64) unfiltered pairs at [qf, qf + n) by their triangles'
// oriented boxes against the owning lanes' thresholds; the survivors are
// appended to the filtered front and the rest of the queue closes up.
__device__ __forceinline__ void q_filter(WaveQ* __restrict__ Q, int n, int& qf, int& qn,
                                         const float4* __restrict__ tobbs, v3 p, float thr, int lane) {
  const int32_t t = lane < n ? Q->qt[qf + lane] : 0;
  const int L = lane < n ? (int)Q->ql[qf + lane] : 0;
  const v3 pl = mk(__shfl(p.x, L), __shfl(p.y, L), __shfl(p.z, L));
  const float tl = __shfl(thr, L);
  bool keep = false;
  if (lane < n) {
    const float4* o = tobbs + 3 * (int64_t)t;
    keep = obb_d2(pl, o[0], o[1], o[2], obb_eq(pl)) <= tl;
  }
  const uint64_t m = __ballot(keep);
  wave_sync();
  if (keep) {
    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    Q->qt[qf + below] = t;
    Q->ql[qf + below] = (uint8_t)L;
  }
  const int s = __popcll(m);
  const int rest = qn - qf - n;
  q_move(Q, qf + n, qf + s, rest, lane);
  qf += s;
  qn = qf + rest;
}

// Tests the first n (1..64) queued pairs exactly and removes them from the
// queue (qn entries before, qn - n after).
__device__ __forceinline__ void q_flush(WaveQ* __restrict__ Q, int n, int qn, const Tri48* __restrict__ tris, v3 p,
                                        int lane, bool valid, float slack, float& best, int32_t& best_id,
                                        int32_t& best_tri, float& thr) {
  const unsigned long long mine = best_key(best, best_id);
  Q->key[lane] = mine;
  const int32_t t = lane < n ? Q->qt[lane] : 0;
  const int L = lane < n ? (int)Q->ql[lane] : 0;
  wave_sync();
  const float px = __shfl(p.x, L), py = __shfl(p.y, L), pz = __shfl(p.z, L);
  unsigned long long k = ~0ull;
  if (lane < n) {
    v3 ta, tb, tc;
    int32_t id;
    load_tri(tris, t, ta, tb, tc, id);
    k = best_key(tri_dist2(ta, tb, tc, mk(px, py, pz)), id);
    atomicMin(&Q->key[L], k);
  }
  wave_sync();
  if (lane < n && Q->key[L] == k) Q->tri[L] = t;
  q_move(Q, n, 0, qn - n, lane);  // the untested tail to the front
  const unsigned long long nk = Q->key[lane];
  if (nk != mine) {
    best = __uint_as_float((uint32_t)(nk >> 32));
    best_id = (int32_t)(uint32_t)(nk & 0xffffffffull);
    best_tri = Q->tri[lane];
    thr = valid ? prune_thr(best, slack) : -1.0f;
  }
  wave_sync();
}

// ------------------------------------------------------------ wave-coherent walk
// Per-lane bounds of the RING most recent stack entries (RING a power of
// two, or 0 for none), indexed by the wave-uniform stack position modulo
// RING.  4 entries: the closest-point walk 12.62 -> 12.12 ms on the 4.1M-face
// union.  The batched walk keeps 2, its kernel held to 80 VGPRs (6
// wavefronts per SIMD, as without the ring): 43.2 -> 41.6 ms on 256 meshes;
// unheld, 4 entries took it to 86 VGPRs, 5 wavefronts and 45.6 ms.
constexpr int kRingUdf = 4;
constexpr int kRingBatch = 2;
template <int RING>
struct WalkRing {
  static_assert(RING >= 0 && (RING & (RING - 1)) == 0, "ring size: 0 or a power of two");
  static constexpr int kMask = RING > 1 ? RING - 1 : 0;
  typedef float V __attribute__((ext_vector_type(RING > 0 ? RING : 1)));
};
// a vector value, not a struct member: dynamic indices into it become selects
// over registers instead of a private-memory array
template <int RING>
__device__ __forceinline__ void ring_put(typename WalkRing<RING>::V& r, int e, float d) {
  r[e & WalkRing<RING>::kMask] = d;
}
template <int RING>
__device__ __forceinline__ float ring_get(const typename WalkRing<RING>::V& r, int e) {
  return r[e & WalkRing<RING>::kMask];
}

// The wave-coherent packet walk of one wavefront (see the file comment):
// each lane's best squared distance / face enter as the seed state and leave
// as the result.  stk: this wave's LDS stack.  n_*: work counters (STATS).
template <bool STATS, bool QUEUE, int RING>
__device__ __forceinline__ void wave_walk(const Node64* __restrict__ nodes, const Tri48* __restrict__ tris,
                                          const TriBox32* __restrict__ boxes, const Obb96* __restrict__ obbs,
                                          const float4* __restrict__ tobbs, v3 p, bool valid, float slack,
                                          float4 (*stk)[2], WaveQ* __restrict__ Q, int lane,
                                          float& best, int32_t& best_id, int32_t& best_tri, uint32_t& n_node,
                                          uint32_t& n_leaf_tri, uint32_t& n_test, uint32_t& n_lane) {
  float thr = valid ? prune_thr(best, slack) : -1.0f;  // invalid lanes accept nothing
  const float eq = obb_eq(p);
  int qn = 0, qf = 0;  // queued pairs, of which the first qf are filtered (QUEUE)
  int sp = 0;
  int32_t cur = 0;
  // Each lane's bound of the last RING pushed children, kept in registers: a pop of
  // entry e reads it instead of recomputing the box and oriented-box bounds, which give
  // the same value (same box, same point, same operations).  Entries [ring_lo, sp) are
  // held; a push at e overwrites entry e - RING.
  typename WalkRing<RING>::V ring;
  int ring_lo = RING > 0 ? 0 : kWaveStack;
  while (true) {
    if (cur >= 0) {
      if (STATS) ++n_node;
      const bool obb_lds = QUEUE && obbs;
      if (obb_lds) {
        // start copying the node's oriented boxes (6 x 16 B, lanes 0-5) into LDS so that they
        // arrive with the node instead of one round trip after its box test; the wait first
        // retires the previous copy, whose slot this one overwrites
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
        if (lane < 6)
          __builtin_amdgcn_global_load_lds(obbs[cur].c + lane, (__attribute__((address_space(3))) void*)Q->ob, 16, 0,
                                           0);
      }
      const Node64 nd = nodes[cur];
      float dl = box_d2(p, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
      float dr = box_d2(p, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
      bool wl = dl <= thr, wr = dr <= thr;
      uint64_t bl = __ballot(wl), br = __ballot(wr);
      if (obbs && (bl | br)) {  // tighten by the children's oriented boxes (those some lane still wants)
        const float4* ob = obbs[cur].c;
        if (obb_lds) {
          __builtin_amdgcn_s_waitcnt(kWaitVm0);
          wave_sync();
          ob = Q->ob;
        }
        if (bl) {
          dl = fmaxf(dl, obb_d2(p, ob[0], ob[1], ob[2], eq));
          wl = dl <= thr;
          bl = __ballot(wl);
        }
        if (br) {
          dr = fmaxf(dr, obb_d2(p, ob[3], ob[4], ob[5], eq));
          wr = dr <= thr;
          br = __ballot(wr);
        }
      }
      if (bl && br) {
        int nl = __popcll(__ballot((wl || wr) && dl <= dr));
        int nr = __popcll(bl | br) - nl;
        bool left_first = nl >= nr;
        int32_t near_ref = left_first ? nd.d.x : nd.d.y;
        int32_t far_ref = left_first ? nd.d.y : nd.d.x;
        if (sp < kWaveStack) {
          if constexpr (RING > 0) {
            ring_put<RING>(ring, sp, left_first ? dr : dl);
            ring_lo = max(ring_lo, sp + 1 - RING);
          }
          if (lane == 0) {
            // w of the second half: the far child as node * 2 + side, for its oriented box
            if (left_first) {
              stk[sp][0] = make_float4(nd.b.z, nd.b.w, nd.c.x, __int_as_float(far_ref));
              stk[sp][1] = make_float4(nd.c.y, nd.c.z, nd.c.w, __int_as_float(cur * 2 + 1));
            } else {
              stk[sp][0] = make_float4(nd.a.x, nd.a.y, nd.a.z, __int_as_float(far_ref));
              stk[sp][1] = make_float4(nd.a.w, nd.b.x, nd.b.y, __int_as_float(cur * 2));
            }
          }
          ++sp;
        }
        cur = uni(near_ref);
        continue;
      }
      if (bl) { cur = uni(nd.d.x); continue; }
      if (br) { cur = uni(nd.d.y); continue; }
    } else {
      const int32_t s0 = leaf_start(cur), n = leaf_count(cur);
      // the leaf's triangle boxes by scalar loads (the address is the same on every lane):
      // the bound test takes them as SGPR operands, with no readlane per coordinate
      typedef const __attribute__((address_space(4))) float CFloat;
      CFloat* cb = (CFloat*)boxes;
      for (int32_t t = s0; t < s0 + n; ++t) {
        v3 ta, tb, tc;
        int32_t id = 0;
        bool need;
        if (QUEUE) {  // bound test from the stored box; vertices only in the flush
          CFloat* b = cb + 8 * (int64_t)t;
          need = box_d2(p, b[0], b[1], b[2], b[4], b[5], b[6]) <= thr;
        } else {
          load_tri(tris, t, ta, tb, tc, id);
          need = tri_box_d2(ta, tb, tc, tri_pad_of(tris, t), p) <= thr;
        }
        if (STATS) ++n_leaf_tri;
        if (STATS) n_lane += need ? 1u : 0u;
        const uint64_t m = __ballot(need);
        if (!m) continue;
        if (QUEUE) {
          if (need) {
            const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            Q->qt[qn + below] = t;
            Q->ql[qn + below] = (uint8_t)lane;
          }
          qn += __popcll(m);
          if (!tobbs) qf = qn;
          if (qn - qf >= 64) {
            wave_sync();
            q_filter(Q, 64, qf, qn, tobbs, p, thr, lane);
          }
          if (qf >= 64) {
            wave_sync();
            if (STATS) ++n_test;
            q_flush(Q, 64, qn, tris, p, lane, valid, slack, best, best_id, best_tri, thr);
            qn -= 64;
            qf -= 64;
          }
          continue;
        }
        if (STATS) ++n_test;
        float d2 = tri_dist2(ta, tb, tc, p);
        if (d2 < best || (d2 == best && id < best_id)) {
          best = d2;
          best_id = id;
          best_tri = t;
          thr = valid ? prune_thr(best, slack) : -1.0f;
        }
      }
    }
    bool found = false;
    while (sp > 0) {
      --sp;
      float d;
      if (RING > 0 && sp >= ring_lo) {
        d = ring_get<RING>(ring, sp);
      } else {
        const float4 e0 = stk[sp][0], e1 = stk[sp][1];
        d = box_d2(p, e0.x, e0.y, e0.z, e1.x, e1.y, e1.z);
        if (obbs && __ballot(d <= thr)) {
          const int32_t code = uni(__float_as_int(e1.w));
          const float4* c = obbs[code >> 1].c + 3 * (code & 1);
          d = fmaxf(d, obb_d2(p, c[0], c[1], c[2], eq));
        }
        if (RING > 0) ring_lo = sp;
      }
      if (__ballot(d <= thr)) {
        cur = uni(__float_as_int(stk[sp][0].w));
        found = true;
        break;
      }
    }
    if (!found) {
      if (QUEUE && qn > 0) {  // drain, then re-check the stack with the tightened thresholds
        wave_sync();
        if (qn > qf) q_filter(Q, qn - qf, qf, qn, tobbs, p, thr, lane);
        while (qn > 0) {
          if (STATS) ++n_test;
          const int n = qn < 64 ? qn : 64;
          q_flush(Q, n, qn, tris, p, lane, valid, slack, best, best_id, best_tri, thr);
          qn -= n;
        }
        qf = 0;
      }
      break;
    }
  }
}

// ------------------------------------------------------------ grid bounds
// A grid of exact distances G and closest faces GF at lo + (i, j, k) h
// (float: lo + (float)i * h; i, j, k in [0, R]), the mesh V, F for the faces.
struct GridBound {  // also the 64-byte record of a3d_batch_udf's bounds array (a3d_grid_bound)
  const float* G;
  const int32_t* GF;
  const float* V;
  const int32_t* F;
  int R;
  float lo[3];
  float h;
  float margin;
  int32_t pad[2];
};
static_assert(sizeof(GridBound) == 64, "grid bound record is 64 bytes");

// Upper bound on the distance from q: the minimum of (a) over the 8 corners
// c of q's cell (clamped into the grid), G(c) + |q - c| (the distance is
// 1-Lipschitz), widened by 2^-12 relative and `margin` absolute for FP32
// rounding, and (b) with GF, the walk's own distance from q to the nearest
// corner's face (a distance some face attains).
__device__ __forceinline__ float grid_bound(const GridBound& g, const float q[3]) {
  const int R = g.R;
  int c0[3];
  for (int d = 0; d < 3; ++d) {
    const float f = floorf((q[d] - g.lo[d]) / g.h);
    c0[d] = f >= 0.0f ? (f < (float)(R - 1) ? (int)f : R - 1) : 0;  // NaN -> 0
  }
  float best = __int_as_float(0x7f800000), best_f2 = best, near_e2 = best;
  int64_t near_g = 0;
  for (int k = 0; k < 8; ++k) {
    const int c[3] = {c0[0] + (k & 1), c0[1] + ((k >> 1) & 1), c0[2] + ((k >> 2) & 1)};
    float e2 = 0.0f;
    for (int d = 0; d < 3; ++d) {
      const float e = q[d] - (g.lo[d] + (float)c[d] * g.h);  // the grid point as the caller placed it
      e2 += e * e;
    }
    const int64_t gi = ((int64_t)c[0] * (R + 1) + c[1]) * (R + 1) + c[2];
    best = fminf(best, g.G[gi] + sqrtf(e2));
    if (e2 < near_e2) {
      near_e2 = e2;
      near_g = gi;
    }
  }
  if (g.GF) {
    const int32_t f = g.GF[near_g];
    if (f >= 0) {
      const int32_t* fv = g.F + 3 * (int64_t)f;
      const float* va = g.V + 3 * (int64_t)fv[0];
      const float* vb = g.V + 3 * (int64_t)fv[1];
      const float* vc = g.V + 3 * (int64_t)fv[2];
      best_f2 = tri_dist2(mk(va[0], va[1], va[2]), mk(vb[0], vb[1], vb[2]), mk(vc[0], vc[1], vc[2]),
                          mk(q[0], q[1], q[2]));
    }
  }
  // x (1 + 2^-12): FP32 rounding of the sum and the grid; a face's own
  // distance needs only the rounding of the square root
  return fminf(best * 1.000244140625f + g.margin, sqrtf(best_f2) * 1.00000095367431640625f);
}

// seed (optional, indexed like the outputs): an upper bound on each query's
// distance; the walk starts with it as the best distance, so it never enters
// a box farther than the bound.  A query with no face within its bound gets
// face -1 and dist = its seed.  gb.G (when seed is null): the bound from a
// grid, computed here (grid_bound).
template <bool STATS, bool QUEUE>
__global__ __launch_bounds__(kBlkW) void k_udf_wave(const void* __restrict__ blob, const float* __restrict__ P,
                                                    const uint32_t* __restrict__ perm, int64_t K,
                                                    const float* __restrict__ seed,
                                                    float* __restrict__ dist, int32_t* __restrict__ face,
                                                    float* __restrict__ closest, float* __restrict__ uvw,
                                                    unsigned long long* __restrict__ counters, bool use_obbs,
                                                    GridBound gb, uint32_t* __restrict__ work) {
  __shared__ float4 s_stk[kWaveStack][2];  // (lo.xyz, ref bits), (hi.xyz, obb code)
  __shared__ WaveQ s_q[1];
  const int lane = threadIdx.x;
  const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlkW + threadIdx.x;
  const bool valid = i < K;
  if (!__any(valid)) return;
  const int64_t q = valid ? (perm ? (int64_t)perm[i] : i) : 0;
  const BvhHeader* H = header(blob);
  const Node64* __restrict__ nodes = nodes_of(blob);
  const Tri48* __restrict__ tris = tris_of(blob);
  const v3 p = valid ? mk(P[q * 3], P[q * 3 + 1], P[q * 3 + 2]) : mk(0.0f, 0.0f, 0.0f);
  const float slack = query_slack(p, H->coord_max);

  float best = 1e30f;
  if (valid && (seed || gb.G)) {
    const float q3[3] = {p.x, p.y, p.z};
    const float sd = seed ? seed[q] : grid_bound(gb, q3);
    if (sd >= 0.0f && sd < 1e15f) best = fmaxf(sd * sd * (1.0f + 4.8e-7f), 1e-30f);
  }
  int32_t best_id = 0x7fffffff, best_tri = -1;
  uint32_t n_node = 0, n_leaf_tri = 0, n_test = 0, n_lane = 0;
  wave_walk<STATS, QUEUE, kRingUdf>(nodes, tris, boxes_of(blob), use_obbs ? obbs_of(blob) : nullptr,
                          use_obbs ? tobbs_of(blob) : nullptr, p, valid, slack,
                          s_stk, QUEUE ? &s_q[0] : nullptr, lane,
                          best, best_id,
                          best_tri, n_node, n_leaf_tri, n_test, n_lane);
  if (STATS) {
    if (work && valid) work[q] = n_lane;
    if (!valid) n_lane = 0;
    for (int o = 32; o > 0; o >>= 1) n_lane += __shfl_xor(n_lane, o);
    if (counters && lane == 0) {
      atomicAdd(&counters[0], 1ull);
      atomicAdd(&counters[1], (unsigned long long)n_node);
      atomicAdd(&counters[2], (unsigned long long)n_leaf_tri);
      atomicAdd(&counters[3], (unsigned long long)n_test);
      atomicAdd(&counters[4], (unsigned long long)n_lane);
    }
  }
  if (valid) {
    if (seed && best_tri < 0) best = seed[q] * seed[q];
    write_result(q, p, best, best_id, best_tri, tris, dist, face, closest, uvw);
  }
}


// One packet of the batched walk: lane's query q (if valid) on `blob`.
// gb (nullable): the mesh's grid bound (see grid_bound) for each lane's walk.
__device__ __forceinline__ void batch_walk_tile(const void* __restrict__ blob, const float* __restrict__ P, int64_t q,
                                                bool valid, float4 (*stk)[2], WaveQ* __restrict__ Q, int lane,
                                                float* __restrict__ dist,
                                                int32_t* __restrict__ face, float* __restrict__ closest,
                                                float* __restrict__ uvw, const GridBound* __restrict__ gb) {
  const BvhHeader* H = header(blob);
  const Node64* __restrict__ nodes = nodes_of(blob);
  const Tri48* __restrict__ tris = tris_of(blob);
  const v3 p = valid ? mk(P[q * 3], P[q * 3 + 1], P[q * 3 + 2]) : mk(0.0f, 0.0f, 0.0f);
  const float slack = query_slack(p, H->coord_max);
  float best = 1e30f;
  if (gb && valid) {
    const GridBound g = *gb;
    const float q3[3] = {p.x, p.y, p.z};
    const float sd = grid_bound(g, q3);
    if (sd >= 0.0f && sd < 1e15f) best = fmaxf(sd * sd * (1.0f + 4.8e-7f), 1e-30f);
  }
  int32_t best_id = 0x7fffffff, best_tri = -1;
  uint32_t n0 = 0, n1 = 0, n2 = 0, n3 = 0;
  wave_walk<false, true, kRingBatch>(nodes, tris, boxes_of(blob), obbs_of(blob), tobbs_of(blob), p, valid, slack,
                                  stk, Q, lane,
                         best, best_id,
                         best_tri, n0, n1, n2, n3);
  if (valid) write_result(q, p, best, best_id, best_tri, tris, dist, face, closest, uvw);
}

// ------------------------------------------------------------ per-lane walk
// redo (the face output, or null): walk only the queries whose face is -1
// there, without a seed (the second pass of a seeded walk).
__global__ __launch_bounds__(kBlk) void k_udf_lane(const void* __restrict__ blob, const float* __restrict__ P,
                                                    const uint32_t* __restrict__ perm, int64_t K,
                                                    const float* __restrict__ seed, float* __restrict__ dist,
                                                    int32_t* face,  // no __restrict__: redo may alias it
                                                    float* __restrict__ closest, float* __restrict__ uvw,
                                                    bool use_obbs, const int32_t* redo, GridBound gb) {
  __shared__ int32_t s_ref[kLaneLds * kBlk];
  __shared__ float s_d2[kLaneLds * kBlk];
  const int tid = threadIdx.x;
  const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlk + tid;
  if (i >= K) return;
  const int64_t q = perm ? (int64_t)perm[i] : i;
  if (redo && redo[q] >= 0) return;
  const BvhHeader* H = header(blob);
  const Node64* __restrict__ nodes = nodes_of(blob);
  const Tri48* __restrict__ tris = tris_of(blob);
  const v3 p = mk(P[q * 3], P[q * 3 + 1], P[q * 3 + 2]);
  const float slack = query_slack(p, H->coord_max);
  const Obb96* __restrict__ obbs = use_obbs ? obbs_of(blob) : nullptr;
  const float eq = obb_eq(p);

  float best = 1e30f;
  if (seed || gb.G) {
    const float q3[3] = {p.x, p.y, p.z};
    const float sd = seed ? seed[q] : grid_bound(gb, q3);
    if (sd >= 0.0f && sd < 1e15f) best = fmaxf(sd * sd * (1.0f + 4.8e-7f), 1e-30f);
  }
  int32_t best_id = 0x7fffffff, best_tri = -1;
  float thr = prune_thr(best, slack);
  int32_t ovf_ref[kLaneOvf];
  float ovf_d2[kLaneOvf];
  int sp = 0;
  int32_t cur = 0;
  while (true) {
    if (cur >= 0) {
      const Node64* nd = nodes + cur;
      float4 a = nd->a, b = nd->b, c = nd->c;
      int4 d = nd->d;
      float dl = box_d2(p, a.x, a.y, a.z, a.w, b.x, b.y);
      float dr = box_d2(p, b.z, b.w, c.x, c.y, c.z, c.w);
      bool tl = dl <= thr, tr = dr <= thr;
      if (obbs && (tl || tr)) {
        const Obb96* ob = obbs + cur;
        dl = fmaxf(dl, obb_d2(p, ob->c[0], ob->c[1], ob->c[2], eq));
        dr = fmaxf(dr, obb_d2(p, ob->c[3], ob->c[4], ob->c[5], eq));
        tl = dl <= thr;
        tr = dr <= thr;
      }
      if (tl && tr) {
        int32_t near_ref = d.x, far_ref = d.y;
        float far_d = dr;
        if (dr < dl) { near_ref = d.y; far_ref = d.x; far_d = dl; }
        if (sp < kLaneLds) {
          s_ref[sp * kBlk + tid] = far_ref;
          s_d2[sp * kBlk + tid] = far_d;
        } else if (sp < kStackTotal) {
          ovf_ref[sp - kLaneLds] = far_ref;
          ovf_d2[sp - kLaneLds] = far_d;
        }
        sp = sp < kStackTotal ? sp + 1 : sp;
        cur = near_ref;
        continue;
      }
      if (tl) { cur = d.x; continue; }
      if (tr) { cur = d.y; continue; }
    } else {
      const int32_t s0 = leaf_start(cur), n = leaf_count(cur);
      for (int32_t t = s0; t < s0 + n; ++t) {
        v3 ta, tb, tc;
        int32_t id;
        load_tri(tris, t, ta, tb, tc, id);
        if (tri_box_d2(ta, tb, tc, tri_pad_of(tris, t), p) > thr) continue;
        float d2 = tri_dist2(ta, tb, tc, p);
        if (d2 < best || (d2 == best && id < best_id)) {
          best = d2;
          best_id = id;
          best_tri = t;
          thr = prune_thr(best, slack);
        }
      }
    }
    bool found = false;
    while (sp > 0) {
      --sp;
      int32_t r;
      float dd;
      if (sp < kLaneLds) {
        r = s_ref[sp * kBlk + tid];
        dd = s_d2[sp * kBlk + tid];
      } else {
        r = ovf_ref[sp - kLaneLds];
        dd = ovf_d2[sp - kLaneLds];
      }
      if (dd <= thr) { cur = r; found = true; break; }
    }
    if (!found) break;
  }
  if (seed && best_tri < 0) best = seed[q] * seed[q];
  write_result(q, p, best, best_id, best_tri, tris, dist, face, closest, uvw);
}

// ------------------------------------------------------------ split per-lane walk
// For few queries (flag A3D_UDF_SPLIT): 4 lanes per query, lane s of a group
// walking the subtree under the root's s-th grandchild (or the root's leaf
// child), so 4x as many wavefronts hide the walk's latency.  Each lane first
// descends the whole tree greedily (nearer child, no stack) and tests that
// leaf exactly, which bounds the query before the sub-walks start.  The
// lanes of a query publish their best squared distances in LDS and prune
// with the group's minimum: any lane's best is an upper bound on the
// query's distance, so the pruning stays conservative.  The result is the group's
// minimum (distance bits, face id) key: the same closest face, ties to the
// lowest index, as one walk over the whole tree.
__global__ __launch_bounds__(kBlk) void k_udf_split(const void* __restrict__ blob, const float* __restrict__ P,
                                                     const uint32_t* __restrict__ perm, int64_t K,
                                                     float* __restrict__ dist, int32_t* __restrict__ face,
                                                     float* __restrict__ closest, float* __restrict__ uvw,
                                                     bool use_obbs) {
  __shared__ int32_t s_ref[kLaneLds * kBlk];
  __shared__ float s_d2[kLaneLds * kBlk];
  __shared__ float s_best[kBlk];  // lane 4j + s: sub-walk s's best of the block's query j
  const int tid = threadIdx.x;
  const int sub = tid & 3;
  const int64_t i = ((int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlk + tid) >> 2;
  const bool valid = i < K;
  const int64_t q = valid ? (perm ? (int64_t)perm[i] : i) : 0;
  const BvhHeader* H = header(blob);
  const Node64* __restrict__ nodes = nodes_of(blob);
  const Tri48* __restrict__ tris = tris_of(blob);
  const v3 p = valid ? mk(P[q * 3], P[q * 3 + 1], P[q * 3 + 2]) : mk(0.0f, 0.0f, 0.0f);
  const float slack = query_slack(p, H->coord_max);
  const Obb96* __restrict__ obbs = use_obbs ? obbs_of(blob) : nullptr;
  const float eq = obb_eq(p);
  volatile float* grp = s_best + (tid & ~3);
  grp[sub] = 1e30f;
  wave_sync();

  float best = 1e30f;
  int32_t best_id = 0x7fffffff, best_tri = -1;
  float thr = valid ? prune_thr(best, slack) : -1.0f;
  if (valid) {  // greedy descent of the whole tree: an exact upper bound for every sub-walk
    int32_t g = 0;
    while (g >= 0) {
      const Node64 nd = nodes[g];
      const float dl = box_d2(p, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
      const float dr = box_d2(p, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
      g = dl <= dr ? nd.d.x : nd.d.y;
      if (!(fminf(dl, dr) <= thr)) { g = 0; break; }
    }
    if (g < 0) {
      const int32_t s0 = leaf_start(g), n = leaf_count(g);
      for (int32_t t = s0; t < s0 + n; ++t) {
        v3 ta, tb, tc;
        int32_t id;
        load_tri(tris, t, ta, tb, tc, id);
        const float d2 = tri_dist2(ta, tb, tc, p);
        if (d2 < best || (d2 == best && id < best_id)) {
          best = d2;
          best_id = id;
          best_tri = t;
        }
      }
      thr = prune_thr(best, slack);
      grp[sub] = best;
    }
  }
  // start: child (sub >> 1) of the root; if internal, its child (sub & 1)
  const Node64 root = nodes[0];
  int32_t cur = (sub >> 1) ? root.d.y : root.d.x;
  float start_d = (sub >> 1) ? box_d2(p, root.b.z, root.b.w, root.c.x, root.c.y, root.c.z, root.c.w)
                             : box_d2(p, root.a.x, root.a.y, root.a.z, root.a.w, root.b.x, root.b.y);
  bool active = valid;
  if (cur >= 0) {
    const Node64 c = nodes[cur];
    start_d = (sub & 1) ? box_d2(p, c.b.z, c.b.w, c.c.x, c.c.y, c.c.z, c.c.w)
                        : box_d2(p, c.a.x, c.a.y, c.a.z, c.a.w, c.b.x, c.b.y);
    cur = (sub & 1) ? c.d.y : c.d.x;
  } else if (sub & 1) {
    active = false;  // a leaf child of the root belongs to the even lane
  }
  if (!(start_d <= thr)) active = false;  // empty boxes (a mesh without faces) included

  int32_t ovf_ref[kLaneOvf];
  float ovf_d2[kLaneOvf];
  int sp = 0;
  while (active) {
    {
      const float g = fminf(fminf(grp[0], grp[1]), fminf(grp[2], grp[3]));
      thr = fminf(thr, prune_thr(g, slack));
    }
    if (cur >= 0) {
      const Node64* nd = nodes + cur;
      float4 a = nd->a, b = nd->b, c = nd->c;
      int4 d = nd->d;
      float dl = box_d2(p, a.x, a.y, a.z, a.w, b.x, b.y);
      float dr = box_d2(p, b.z, b.w, c.x, c.y, c.z, c.w);
      bool tl = dl <= thr, tr = dr <= thr;
      if (obbs && (tl || tr)) {
        const Obb96* ob = obbs + cur;
        dl = fmaxf(dl, obb_d2(p, ob->c[0], ob->c[1], ob->c[2], eq));
        dr = fmaxf(dr, obb_d2(p, ob->c[3], ob->c[4], ob->c[5], eq));
        tl = dl <= thr;
        tr = dr <= thr;
      }
      if (tl && tr) {
        int32_t near_ref = d.x, far_ref = d.y;
        float far_d = dr;
        if (dr < dl) { near_ref = d.y; far_ref = d.x; far_d = dl; }
        if (sp < kLaneLds) {
          s_ref[sp * kBlk + tid] = far_ref;
          s_d2[sp * kBlk + tid] = far_d;
        } else if (sp < kStackTotal) {
          ovf_ref[sp - kLaneLds] = far_ref;
          ovf_d2[sp - kLaneLds] = far_d;
        }
        sp = sp < kStackTotal ? sp + 1 : sp;
        cur = near_ref;
        continue;
      }
      if (tl) { cur = d.x; continue; }
      if (tr) { cur = d.y; continue; }
    } else {
      const int32_t s0 = leaf_start(cur), n = leaf_count(cur);
      for (int32_t t = s0; t < s0 + n; ++t) {
        v3 ta, tb, tc;
        int32_t id;
        load_tri(tris, t, ta, tb, tc, id);
        if (tri_box_d2(ta, tb, tc, tri_pad_of(tris, t), p) > thr) continue;
        float d2 = tri_dist2(ta, tb, tc, p);
        if (d2 < best || (d2 == best && id < best_id)) {
          best = d2;
          best_id = id;
          best_tri = t;
          thr = fminf(thr, prune_thr(best, slack));
          grp[sub] = best;
        }
      }
    }
    bool found = false;
    while (sp > 0) {
      --sp;
      int32_t r;
      float dd;
      if (sp < kLaneLds) {
        r = s_ref[sp * kBlk + tid];
        dd = s_d2[sp * kBlk + tid];
      } else {
        r = ovf_ref[sp - kLaneLds];
        dd = ovf_d2[sp - kLaneLds];
      }
      if (dd <= thr) { cur = r; found = true; break; }
    }
    if (!found) break;
  }
  // the group's minimum key; every lane takes part in the shuffles
  unsigned long long key = best_key(best, best_id);
  int32_t bt = best_tri;
  for (int o = 1; o < 4; o <<= 1) {
    const unsigned long long ok = __shfl_xor(key, o);
    const int32_t obt = __shfl_xor(bt, o);
    if (ok < key) {
      key = ok;
      bt = obt;
    }
  }
  if (valid && sub == 0)
    write_result(q, p, __uint_as_float((uint32_t)(key >> 32)), (int32_t)(uint32_t)(key & 0xffffffffull), bt, tris,
                 dist, face, closest, uvw);
}

// ------------------------------------------------------------ query ordering
__device__ __forceinline__ uint32_t f2ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_qinit(uint32_t* red) {
  int t = threadIdx.x;
  if (t < 6) red[t] = t < 3 ? 0xffffffffu : 0u;
}

__global__ __launch_bounds__(kBlk) void k_qbounds(const float* __restrict__ P, int64_t K, uint32_t* red) {
  float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
  for (int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x; i < K; i += (int64_t)gridDim.x * kBlk) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float v = P[i * 3 + k];
      lo[k] = fminf(lo[k], v);
      hi[k] = fmaxf(hi[k], v);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    for (int o = 32; o > 0; o >>= 1) {
      lo[k] = fminf(lo[k], __shfl_xor(lo[k], o));
      hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], o));
    }
  }
  // one atomic per block and bound: same-address atomics serialise
  __shared__ float s_lo[kWaves][3], s_hi[kWaves][3];
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 3; ++k) {
      s_lo[wv][k] = lo[k];
      s_hi[wv][k] = hi[k];
    }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    float a = s_lo[0][k], b = s_hi[0][k];
    for (int w = 1; w < kWaves; ++w) {
      a = fminf(a, s_lo[w][k]);
      b = fmaxf(b, s_hi[w][k]);
    }
    atomicMin(&red[k], f2ord(a));
    atomicMax(&red[3 + k], f2ord(b));
  }
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__device__ __forceinline__ uint32_t compact3(uint32_t v) {  // every third bit -> low bits
  v &= 0x09249249u;
  v = (v ^ (v >> 2)) & 0x030c30c3u;
  v = (v ^ (v >> 4)) & 0x0300f00fu;
  v = (v ^ (v >> 8)) & 0xff0000ffu;
  v = (v ^ (v >> 16)) & 0x000003ffu;
  return v;
}

__global__ __launch_bounds__(kBlk) void k_qmorton(const float* __restrict__ P, int64_t K,
                                                   const uint32_t* __restrict__ red, uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
  int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i >= K) return;
  uint32_t c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float lo = ord2f(red[k]), ext = ord2f(red[3 + k]) - lo;
    float t = ext > 0.0f ? (P[i * 3 + k] - lo) * (1023.0f / ext) : 0.0f;
    t = fminf(fmaxf(t, 0.0f), 1023.0f);
    c[k] = (uint32_t)t;
  }
  keys[i] = (spread10(c[0]) << 2) | (spread10(c[1]) << 1) | spread10(c[2]);
  vals[i] = (uint32_t)i;
}

// Cost class of each full 64-query packet of the Morton order (one
// wavefront per packet): the largest work of the grid points around the
// packet's queries, each query placed at the centre of its Morton cell
// (decoded from its sorted key: 1/1024 of the queries' box per axis, read
// coalesced instead of gathering the query) and reading the 8 corners of
// its grid cell (clamped as grid_bound clamps).  key = 32 - floor(log2(max
// + 1)), so a stable ascending sort puts the heaviest class first and keeps
// the Morton order inside each class.  The class only orders the walk's
// packets; no result depends on it.
__global__ __launch_bounds__(kBlk) void k_qpacket(const uint32_t* __restrict__ mkey, const uint32_t* __restrict__ red,
                                                  int64_t nfull, const uint32_t* __restrict__ work, int R, float lo0,
                                                  float lo1, float lo2, float h, uint32_t* __restrict__ pkey,
                                                  uint32_t* __restrict__ pval) {
  const int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i >= nfull) return;  // nfull is a multiple of 64: whole wavefronts leave together
  const uint32_t m = mkey[i];
  const uint32_t cm[3] = {compact3(m >> 2), compact3(m >> 1), compact3(m)};
  const float lo[3] = {lo0, lo1, lo2};
  int c0[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float qlo = ord2f(red[d]), ext = ord2f(red[3 + d]) - qlo;
    const float x = qlo + ((float)cm[d] + 0.5f) * (ext * (1.0f / 1023.0f));
    const float f = floorf((x - lo[d]) / h);
    c0[d] = f >= 0.0f ? (f < (float)(R - 1) ? (int)f : R - 1) : 0;  // NaN -> 0
  }
  uint32_t w = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int64_t gi = ((int64_t)(c0[0] + (k & 1)) * (R + 1) + c0[1] + ((k >> 1) & 1)) * (R + 1) + c0[2] + (k >> 2);
    w = max(w, work[gi]);
  }
  for (int o = 32; o > 0; o >>= 1) w = max(w, (uint32_t)__shfl_xor((int)w, o));
  if ((threadIdx.x & 63) == 0) {
    const int64_t j = i >> 6;
    pkey[j] = (uint32_t)(32 - (63 - __builtin_clzll((unsigned long long)w + 1ull)));
    pval[j] = (uint32_t)j;
  }
}

// perm = the packets of pm in the order porder, the trailing partial packet last.
__global__ __launch_bounds__(kBlk) void k_qgather(const uint32_t* __restrict__ pm, const uint32_t* __restrict__ porder,
                                                  int64_t K, int64_t nfull, uint32_t* __restrict__ perm) {
  const int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i >= K) return;
  perm[i] = i < nfull ? pm[(int64_t)porder[i >> 6] * 64 + (i & 63)] : pm[i];
}

// The query order sorts the 30-bit Morton codes (10 bits per axis over the
// queries' box) on their top 24 bits, three 8-bit radix passes instead of
// four: ties, queries within one 4 x 4 x 4 block of the finest cells, keep
// their input order (the sort is stable), which changes nothing for a
// 64-query packet.
constexpr int kMortonLowBits = 6;

struct QWs {
  uint32_t* red;
  uint32_t* k0;
  uint32_t* k1;
  uint32_t* v0;
  uint32_t* pm;  // packet order only: the Morton order, then per packet key/value in and out
  uint32_t* pk0;
  uint32_t* pk1;
  uint32_t* pv0;
  uint32_t* pv1;
  void* cub;
  size_t cub_bytes;
  int64_t total;
};

inline int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

QWs carve_q(void* base, int64_t K, bool packets = false) {
  QWs w{};
  int64_t n = K > 1 ? K : 2;
  char* b = (char*)base;
  int64_t off = 0;
  auto take = [&](int64_t bytes) { char* p = b ? b + off : nullptr; off += al256(bytes); return p; };
  w.red = (uint32_t*)take(64);
  w.k0 = (uint32_t*)take(4 * n);
  w.k1 = (uint32_t*)take(4 * n);
  w.v0 = (uint32_t*)take(4 * n);
  size_t tb = 0;
  (void)rocprim::radix_sort_pairs(nullptr, tb, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (uint32_t*)nullptr, (int)n, kMortonLowBits, 30);
  if (packets) {
    const int64_t np = n / 64 > 1 ? n / 64 : 2;
    w.pm = (uint32_t*)take(4 * n);
    w.pk0 = (uint32_t*)take(4 * np);
    w.pk1 = (uint32_t*)take(4 * np);
    w.pv0 = (uint32_t*)take(4 * np);
    w.pv1 = (uint32_t*)take(4 * np);
    size_t tp = 0;
    (void)rocprim::radix_sort_pairs(nullptr, tp, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                    (uint32_t*)nullptr, (int)np, 0, 6);
    tb = tb > tp ? tb : tp;
  }
  w.cub_bytes = tb;
  w.cub = take((int64_t)tb);
  w.total = off;
  return w;
}

// The Morton order of P into out (uint32[K], K >= 1) from workspace w.
int morton_sort(const QWs& w, const float* P, int64_t K, uint32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_qinit, dim3(1), dim3(64), 0, s, w.red);
  A3D_LAUNCH_CHECK();
  const unsigned g = (unsigned)div_up(K, kBlk);
  hipLaunchKernelGGL(k_qbounds, dim3(g < 1024 ? g : 1024), dim3(kBlk), 0, s, P, K, w.red);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_qmorton, dim3(g), dim3(kBlk), 0, s, P, K, w.red, w.k0, w.v0);
  A3D_LAUNCH_CHECK();
  size_t tb = w.cub_bytes;
  A3D_HIP_CHECK(rocprim::radix_sort_pairs(w.cub, tb, w.k0, w.k1, w.v0, out, (int)K, kMortonLowBits, 30, s));
  return A3D_OK;
}

inline unsigned grid_for(int64_t n) { return (unsigned)div_up(n > 0 ? n : 1, kBlk); }
inline unsigned grid_for_w(int64_t n) { return (unsigned)div_up(n > 0 ? n : 1, kBlkW); }

}  // namespace
}  // namespace a3d

using namespace a3d;

extern "C" int a3d_query_order_ws_bytes(int64_t K, int64_t* bytes) {
  if (K < 0 || K > 0x7fffffff || !bytes) return fail(A3D_ERR_ARG, "a3d_query_order_ws_bytes: K=%lld not in [0, 2^31)", (long long)K);
  *bytes = carve_q(nullptr, K).total;
  return A3D_OK;
}

extern "C" int a3d_query_order(const float* P, int64_t K, uint32_t* perm, void* ws, int64_t ws_bytes,
                               a3d_stream_t stream) {
  int64_t need = 0;
  if (int rc = a3d_query_order_ws_bytes(K, &need)) return rc;
  if (K == 0) return A3D_OK;
  if (!P || !perm || !ws) return fail(A3D_ERR_ARG, "a3d_query_order: null argument");
  if (ws_bytes < need) return fail(A3D_ERR_CAPACITY, "a3d_query_order: workspace %lld < %lld bytes", (long long)ws_bytes, (long long)need);
  if ((uintptr_t)ws & 255) return fail(A3D_ERR_ARG, "a3d_query_order: workspace must be 256-byte aligned");
  return morton_sort(carve_q(ws, K), P, K, perm, (hipStream_t)stream);
}

extern "C" int a3d_query_order_work_ws_bytes(int64_t K, int64_t* bytes) {
  if (K < 0 || K > 0x7fffffff || !bytes)
    return fail(A3D_ERR_ARG, "a3d_query_order_work_ws_bytes: K=%lld not in [0, 2^31)", (long long)K);
  *bytes = carve_q(nullptr, K, true).total;
  return A3D_OK;
}

// The Morton order of the queries with its 64-query packets regrouped by
// estimated cost, heaviest first (longest-processing-time first: a packet
// that walks far longer than the rest, started late, would end the walk
// late), the Morton order kept within each cost class.  work: uint32 per
// grid point (a3d_bvh_udf_work's output for the points of a3d_grid_points),
// on the grid lo3 + (i, j, k) h, i, j, k in [0, R].
extern "C" int a3d_query_order_work(const float* P, int64_t K, const uint32_t* work, int R, const float* lo3, float h,
                                    uint32_t* perm, void* ws, int64_t ws_bytes, a3d_stream_t stream) {
  int64_t need = 0;
  if (int rc = a3d_query_order_work_ws_bytes(K, &need)) return rc;
  if (R < 1 || R > 2048 || !(h > 0.0f) || !lo3)
    return fail(A3D_ERR_ARG, "a3d_query_order_work: bad grid (R=%d, h=%g)", R, (double)h);
  if (K == 0) return A3D_OK;
  if (!P || !perm || !ws || !work) return fail(A3D_ERR_ARG, "a3d_query_order_work: null argument");
  if (ws_bytes < need)
    return fail(A3D_ERR_CAPACITY, "a3d_query_order_work: workspace %lld < %lld bytes", (long long)ws_bytes, (long long)need);
  if ((uintptr_t)ws & 255) return fail(A3D_ERR_ARG, "a3d_query_order_work: workspace must be 256-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const QWs w = carve_q(ws, K, true);
  const int64_t np = K / 64, nfull = np * 64;
  if (np < 2) return morton_sort(w, P, K, perm, s);
  if (int rc = morton_sort(w, P, K, w.pm, s)) return rc;
  hipLaunchKernelGGL(k_qpacket, dim3(grid_for(nfull)), dim3(kBlk), 0, s, w.k1, w.red, nfull, work, R, lo3[0],
                     lo3[1], lo3[2], h, w.pk0, w.pv0);
  A3D_LAUNCH_CHECK();
  size_t tb = w.cub_bytes;
  A3D_HIP_CHECK(rocprim::radix_sort_pairs(w.cub, tb, w.pk0, w.pk1, w.pv0, w.pv1, (int)np, 0, 6, s));
  hipLaunchKernelGGL(k_qgather, dim3(grid_for(K)), dim3(kBlk), 0, s, w.pm, w.pv1, K, nfull, perm);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

// ------------------------------------------------------------ allocation-free query order
// a3d_query_order with its workspace in a file-scope device array instead
// of a caller buffer, for callers that cannot allocate.  Each GPU holds its
// own copy of the array (kNoAllocScratch bytes) until the process ends; its
// device address is looked up once per device with the device current, and
// host code uses only that address (the array's name is a host-side stand-in).
namespace {
constexpr int64_t kNoAllocScratch = int64_t(48) << 20;  // 2^21 queries need ~42 MB (keys, values, sort state)
constexpr int64_t kNoAllocMaxQueries = int64_t(1) << 21;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kNoAllocScratch];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_ptr(void** out) {
  int dev = 0;
  A3D_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(A3D_ERR_INTERNAL, "a3d_query_order_noalloc: device %d >= %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    A3D_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(A3D_ERR_INTERNAL, "a3d_query_order_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return A3D_OK;
}
}  // namespace

extern "C" int a3d_query_order_noalloc(const float* P, int64_t K, uint32_t* perm, a3d_stream_t stream) {
  if (K < 0 || K > kNoAllocMaxQueries)
    return fail(A3D_ERR_CAPACITY, "a3d_query_order_noalloc: K=%lld not in [0, %lld] (fixed scratch of %lld bytes)",
                (long long)K, (long long)kNoAllocMaxQueries, (long long)kNoAllocScratch);
  int64_t need = 0;
  if (int rc = a3d_query_order_ws_bytes(K, &need)) return rc;
  if (need > kNoAllocScratch)
    return fail(A3D_ERR_CAPACITY, "a3d_query_order_noalloc: K=%lld needs %lld bytes of scratch > %lld",
                (long long)K, (long long)need, (long long)kNoAllocScratch);
  if (K == 0) return A3D_OK;
  void* ws = nullptr;
  if (int rc = scratch_ptr(&ws)) return rc;
  return a3d_query_order(P, K, perm, ws, kNoAllocScratch, stream);
}

namespace {
int udf_launch(const void* blob, const float* P, int64_t K, const uint32_t* perm, const float* seed, float* dist,
               int32_t* face, float* closest, float* uvw, int flags, int64_t* counters, a3d_stream_t stream,
               const GridBound* grid = nullptr, uint32_t* work = nullptr) {
  const GridBound gb = grid ? *grid : GridBound{};
  if (!blob) return fail(A3D_ERR_ARG, "a3d_bvh_udf: null BVH blob");
  // the packet walk launches one 64-lane block per 64 queries (grid x < 2^32)
  if (K < 0 || K > (int64_t(1) << 37)) return fail(A3D_ERR_ARG, "a3d_bvh_udf: K=%lld out of range", (long long)K);
  if (perm && K > 0xffffffffll) return fail(A3D_ERR_ARG, "a3d_bvh_udf: perm needs K < 2^32");
  if (flags & ~(A3D_UDF_LANE | A3D_UDF_DIRECT | A3D_UDF_NO_OBBS | A3D_UDF_SPLIT | A3D_UDF_REDO))
    return fail(A3D_ERR_ARG, "a3d_bvh_udf: unknown flags 0x%x", flags);
  if ((flags & A3D_UDF_REDO) && !seed && !grid) return fail(A3D_ERR_ARG, "a3d_bvh_udf: A3D_UDF_REDO needs seeds");
  if (grid && (flags & A3D_UDF_SPLIT)) return fail(A3D_ERR_ARG, "a3d_bvh_udf_grid: no grid bounds in the split walk");
  const bool use_obbs = !(flags & A3D_UDF_NO_OBBS);
  if ((counters || work) && (flags & (A3D_UDF_LANE | A3D_UDF_SPLIT)))
    return fail(A3D_ERR_ARG, "a3d_bvh_udf: counters and work counts need the wave walk");
  if ((flags & A3D_UDF_SPLIT) && (seed || (flags & (A3D_UDF_LANE | A3D_UDF_DIRECT))))
    return fail(A3D_ERR_ARG, "a3d_bvh_udf: A3D_UDF_SPLIT takes no seed and no other walk flag");
  if ((flags & A3D_UDF_SPLIT) && K > (int64_t(1) << 36))
    return fail(A3D_ERR_ARG, "a3d_bvh_udf: A3D_UDF_SPLIT needs K <= 2^36");
  if (K == 0) return A3D_OK;
  if (!P || !dist || !face || !closest) return fail(A3D_ERR_ARG, "a3d_bvh_udf: null input/output");
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* c = (unsigned long long*)counters;
  if (flags & A3D_UDF_SPLIT)
    hipLaunchKernelGGL(k_udf_split, dim3(grid_for(4 * K)), dim3(kBlk), 0, s, blob, P, perm, K, dist, face, closest,
                       uvw, use_obbs);
  else if (flags & A3D_UDF_LANE)
    hipLaunchKernelGGL(k_udf_lane, dim3(grid_for(K)), dim3(kBlk), 0, s, blob, P, perm, K, seed, dist, face, closest,
                       uvw, use_obbs, (const int32_t*)nullptr, gb);
  else if (flags & A3D_UDF_DIRECT) {
    if (counters || work)
      hipLaunchKernelGGL((k_udf_wave<true, false>), dim3(grid_for_w(K)), dim3(kBlkW), 0, s, blob, P, perm, K, seed, dist,
                         face, closest, uvw, c, use_obbs, gb, work);
    else
      hipLaunchKernelGGL((k_udf_wave<false, false>), dim3(grid_for_w(K)), dim3(kBlkW), 0, s, blob, P, perm, K, seed, dist,
                         face, closest, uvw, c, use_obbs, gb, (uint32_t*)nullptr);
  } else {
    if (counters || work)
      hipLaunchKernelGGL((k_udf_wave<true, true>), dim3(grid_for_w(K)), dim3(kBlkW), 0, s, blob, P, perm, K, seed, dist,
                         face, closest, uvw, c, use_obbs, gb, work);
    else
      hipLaunchKernelGGL((k_udf_wave<false, true>), dim3(grid_for_w(K)), dim3(kBlkW), 0, s, blob, P, perm, K, seed, dist,
                         face, closest, uvw, c, use_obbs, gb, (uint32_t*)nullptr);
  }
  A3D_LAUNCH_CHECK();
  if (flags & A3D_UDF_REDO) {  // queries whose seed was below their distance: walk them again without one
    hipLaunchKernelGGL(k_udf_lane, dim3(grid_for(K)), dim3(kBlk), 0, s, blob, P, (const uint32_t*)nullptr, K,
                       (const float*)nullptr, dist, face, closest, uvw, use_obbs, (const int32_t*)face, GridBound{});
    A3D_LAUNCH_CHECK();
  }
  return A3D_OK;
}

// Distance bounds from a grid of exact distances (see a3d_udf_grid_seed).
__global__ __launch_bounds__(kBlk) void k_grid_seed(GridBound gb, const float* __restrict__ P,
                                                    const uint32_t* __restrict__ perm, int64_t K,
                                                    float* __restrict__ seed) {
  const int64_t j = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (j >= K) return;
  const int64_t i = perm ? (int64_t)perm[j] : j;  // the walk's order: neighbours share cells and faces
  const float q[3] = {P[i * 3], P[i * 3 + 1], P[i * 3 + 2]};
  seed[i] = grid_bound(gb, q);
}

// The points of a side^3 grid, lo + (float)i * h per axis, at their own
// index (i * side + j) * side + k, and a walk order covering them in 4x4x4
// bricks (one wavefront each), the bricks along a Morton curve.
__global__ __launch_bounds__(kBlk) void k_grid_points(float lo0, float lo1, float lo2, float h, int side,
                                                      float* __restrict__ P, uint32_t* __restrict__ perm) {
  const int64_t n = (int64_t)side * side * side;
  const int64_t i = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (i >= n) return;
  const uint32_t b = (uint32_t)(i >> 6), l = (uint32_t)(i & 63);
  const uint32_t x = compact3(b) * 4 + (l & 3), y = compact3(b >> 1) * 4 + ((l >> 2) & 3),
                 z = compact3(b >> 2) * 4 + (l >> 4);
  const int64_t q = ((int64_t)x * side + y) * side + z;
  perm[i] = (uint32_t)q;
  P[q * 3] = lo0 + (float)x * h;
  P[q * 3 + 1] = lo1 + (float)y * h;
  P[q * 3 + 2] = lo2 + (float)z * h;
}
}  // namespace

// Grid points for the exact-distance grid of a3d_udf_grid_seed /
// a3d_bvh_udf_grid: P float32[side^3 * 3] = lo3 + (i, j, k) h (host lo3;
// float: lo + (float)i * h) at index (i side + j) side + k, and perm
// uint32[side^3], a walk order for a3d_bvh_udf (4x4x4 bricks along a Morton
// curve: no sort needed).  side = 4 * 2^m, 4 <= side <= 1024.
extern "C" int a3d_grid_points(const float* lo3, float h, int side, float* P, uint32_t* perm, a3d_stream_t stream) {
  if (!lo3 || !P || !perm) return fail(A3D_ERR_ARG, "a3d_grid_points: null argument");
  if (side < 4 || side > 1024 || (side & (side - 1)) || !(h > 0.0f))
    return fail(A3D_ERR_ARG, "a3d_grid_points: side=%d must be a power of two in [4, 1024], h=%g > 0", side, (double)h);
  const int64_t n = (int64_t)side * side * side;
  hipLaunchKernelGGL(k_grid_points, dim3(grid_for(n)), dim3(kBlk), 0, (hipStream_t)stream, lo3[0], lo3[1], lo3[2], h,
                     side, P, perm);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

// Upper bounds on the closest-point distance of each query from a grid of
// exact distances: G[(i (R+1) + j) (R+1) + k] = the distance at lo + (i, j,
// k) h (float: lo + (float)i * h), i, j, k in [0, R].  For a query q, the
// cell holding q (clamped into the grid) gives seed = min over its 8 corners
// c of G(c) + |q - c| (the distance is 1-Lipschitz), widened by 2^-12
// relative and `margin` absolute, and, with grid_face (the closest face of
// each grid point; V, F the mesh), the walk's own distance from q to the
// nearest corner's face (a distance some face attains: usually far tighter).
// perm (nullable): the order the walk takes the queries in.  A seed that still falls below a query's
// distance only makes the seeded walk report face -1 for it, which
// A3D_UDF_REDO walks again.
extern "C" int a3d_udf_grid_seed(const float* grid, const int32_t* grid_face, const float* V, const int32_t* F,
                                 int R, const float* lo3, float h, const float* P, const uint32_t* perm, int64_t K,
                                 float margin, float* seed, a3d_stream_t stream) {
  if (K < 0 || K > (int64_t(1) << 38)) return fail(A3D_ERR_ARG, "a3d_udf_grid_seed: K=%lld out of range", (long long)K);
  if (R < 1 || R > 2048 || !(h > 0.0f) || !(margin >= 0.0f) || !lo3)
    return fail(A3D_ERR_ARG, "a3d_udf_grid_seed: bad grid (R=%d, h=%g, margin=%g)", R, (double)h, (double)margin);
  if (K == 0) return A3D_OK;
  if (!grid || !P || !seed || (grid_face && (!V || !F)))
    return fail(A3D_ERR_ARG, "a3d_udf_grid_seed: null argument");
  const GridBound gb{grid, grid_face, V, F, R, {lo3[0], lo3[1], lo3[2]}, h, margin, {0, 0}};
  hipLaunchKernelGGL(k_grid_seed, dim3(grid_for(K)), dim3(kBlk), 0, (hipStream_t)stream, gb, P, perm, K, seed);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_bvh_udf(const void* blob, const float* P, int64_t K, const uint32_t* perm, float* dist,
                           int32_t* face, float* closest, float* uvw, int flags, int64_t* counters,
                           a3d_stream_t stream) {
  return udf_launch(blob, P, K, perm, nullptr, dist, face, closest, uvw, flags, counters, stream);
}

// a3d_bvh_udf (wave walk) that also writes each query's work: the number of
// leaf triangles whose bound admitted the query during its packet's walk.
extern "C" int a3d_bvh_udf_work(const void* blob, const float* P, int64_t K, const uint32_t* perm, float* dist,
                                int32_t* face, float* closest, float* uvw, int flags, uint32_t* work,
                                a3d_stream_t stream) {
  if (!work && K > 0) return fail(A3D_ERR_ARG, "a3d_bvh_udf_work: null work output");
  return udf_launch(blob, P, K, perm, nullptr, dist, face, closest, uvw, flags, nullptr, stream, nullptr, work);
}

extern "C" int a3d_bvh_udf_seeded(const void* blob, const float* P, int64_t K, const uint32_t* perm,
                                  const float* seed, float* dist, int32_t* face, float* closest, float* uvw,
                                  int flags, a3d_stream_t stream) {
  if (!seed && K > 0) return fail(A3D_ERR_ARG, "a3d_bvh_udf_seeded: null seed");
  return udf_launch(blob, P, K, perm, seed, dist, face, closest, uvw, flags, nullptr, stream);
}

extern "C" int a3d_bvh_udf_grid(const void* blob, const float* P, int64_t K, const uint32_t* perm,
                                const float* grid, const int32_t* grid_face, const float* V, const int32_t* F, int R,
                                const float* lo3, float h, float margin, float* dist, int32_t* face, float* closest,
                                float* uvw, int flags, a3d_stream_t stream) {
  if (R < 1 || R > 2048 || !(h > 0.0f) || !(margin >= 0.0f) || !lo3)
    return fail(A3D_ERR_ARG, "a3d_bvh_udf_grid: bad grid (R=%d, h=%g, margin=%g)", R, (double)h, (double)margin);
  if (K > 0 && (!grid || (grid_face && (!V || !F)))) return fail(A3D_ERR_ARG, "a3d_bvh_udf_grid: null grid argument");
  const GridBound gb{grid, grid_face, V, F, R, {lo3[0], lo3[1], lo3[2]}, h, margin, {0, 0}};
  return udf_launch(blob, P, K, perm, nullptr, dist, face, closest, uvw, flags | A3D_UDF_REDO, nullptr, stream, &gb);
}

// ============================================================ batched walk
// Batched closest-point queries over many meshes of uneven size in one
// persistent launch, plus HIP-graph capture helpers.  No counterpart in the
// reference (its MeshBVH answers one mesh per call).
//
// Plan (once per batch): every query gets the key mesh << 30 | its 30-bit
// Morton code in its mesh's box; one radix sort makes the query order
// mesh-major and spatially coherent within each mesh.  The caller splits
// each mesh's sorted range into tiles of at most 64 queries (one wavefront
// packet) and may order the tiles (e.g. most expensive meshes first).
// Run: a persistent grid of waves takes tiles from one atomic counter and
// walks each tile as a packet on its mesh's tree (the same walk and the same
// closest-face definition as a3d_bvh_udf, so results are bit-identical),
// writing results at the queries' positions.  Every launch is sync-free and
// can be captured in a graph; the counter is reset by a kernel, not a memset.
namespace a3d {

namespace {

constexpr int kBlkB = 256;

__device__ __forceinline__ uint32_t spread10b(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ void k_batch_keys(const void* const* __restrict__ blobs, int n_meshes, const float* __restrict__ P,
                             int64_t K, const int64_t* __restrict__ q_off, uint64_t* __restrict__ keys,
                             uint32_t* __restrict__ vals) {
  const int64_t i = blockIdx.x * (int64_t)kBlkB + threadIdx.x;
  if (i >= K) return;
  int lo = 0, hi = n_meshes - 1;  // last mesh with q_off[m] <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (q_off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  const BvhHeader* H = header(blobs[lo]);
  uint32_t c[3];
  for (int k = 0; k < 3; ++k) {
    const float l = H->scene_lo[k], ext = H->scene_hi[k] - l;
    float t = ext > 0.0f ? (P[i * 3 + k] - l) * (1023.0f / ext) : 0.0f;
    t = fminf(fmaxf(t, 0.0f), 1023.0f);
    c[k] = (uint32_t)t;
  }
  keys[i] = ((uint64_t)lo << 30) | (spread10b(c[0]) << 2) | (spread10b(c[1]) << 1) | spread10b(c[2]);
  vals[i] = (uint32_t)i;
}

__global__ void k_batch_reset(unsigned int* counter) {
  if (threadIdx.x == 0) *counter = 0u;
}

// tiles: int64 pairs (start << 8 | count, mesh), start = index into perm.
// bounds (nullable): one GridBound per mesh; redo: walk only the lanes whose
// face is -1 (the second pass after a bounded one), without bounds.
__global__ __attribute__((amdgpu_waves_per_eu(6, 6))) void k_udf_batch(const void* const* __restrict__ blobs,
                                                                          const float* __restrict__ P,
                            const uint32_t* __restrict__ perm, const int64_t* __restrict__ tiles, int64_t n_tiles,
                            unsigned int* __restrict__ counter, float* __restrict__ dist, int32_t* __restrict__ face,
                            float* __restrict__ closest, float* __restrict__ uvw,
                            const GridBound* __restrict__ bounds, bool redo) {
  extern __shared__ float4 s_dyn[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  float4(*stk)[2] = reinterpret_cast<float4(*)[2]>(s_dyn + (size_t)wv * kWaveStack * 2);
  WaveQ* Q = reinterpret_cast<WaveQ*>(s_dyn + (size_t)nw * kWaveStack * 2) + wv;
  // the redo pass finds little to do: tiles split statically (wave w takes w, w + W, ...), no queue
  const int64_t wave_id = (int64_t)blockIdx.x * nw + wv, n_waves = (int64_t)gridDim.x * nw;
  for (int64_t it = 0;; ++it) {
    unsigned int t = 0;
    if (redo) {
      const int64_t tt = wave_id + it * n_waves;
      if (tt >= n_tiles) break;
      t = (unsigned int)tt;
    } else {
      if (lane == 0) t = atomicAdd(counter, 1u);
      t = __builtin_amdgcn_readfirstlane(t);
      if ((int64_t)t >= n_tiles) break;
    }
    const int64_t w0 = tiles[2 * (int64_t)t], mesh = tiles[2 * (int64_t)t + 1];
    const int64_t start = w0 >> 8;
    const int count = (int)(w0 & 0xff);
    bool valid = lane < count;
    const int64_t q = valid ? (int64_t)perm[start + lane] : 0;
    if (redo) {
      valid = valid && face[q] < 0;
      if (!__any(valid)) continue;
    }
    batch_walk_tile(blobs[mesh], P, q, valid, stk, Q, lane, dist, face, closest, uvw,
                    (bounds && !redo) ? bounds + mesh : nullptr);
  }
}

size_t sort_bytes(int64_t K) {
  size_t t = 0, t32 = 0;
  (void)rocprim::radix_sort_pairs(nullptr, t, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint32_t*)nullptr,
                                           (uint32_t*)nullptr, (int)(K > 1 ? K : 2), 0, 40);
  (void)rocprim::radix_sort_pairs(nullptr, t32, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                  (uint32_t*)nullptr, (int)(K > 1 ? K : 2), 0, 32);
  return t > t32 ? t : t32;
}

// Up to 256 meshes the batch key fits 32 bits: mesh << 24 | the top 24 of the
// 30 Morton bits (8 per axis over the mesh's box), which halves the sort's
// key traffic and saves a radix pass against the 64-bit key.
__global__ void k_batch_keys32(const uint64_t* __restrict__ k64, int64_t K, uint32_t* __restrict__ k32) {
  const int64_t i = blockIdx.x * (int64_t)kBlkB + threadIdx.x;
  if (i >= K) return;
  const uint64_t k = k64[i];
  k32[i] = (uint32_t)((k >> 30) << 24) | (uint32_t)((k & 0x3fffffffull) >> 6);
}

}  // namespace
}  // namespace a3d

using namespace a3d;

extern "C" int a3d_batch_order_ws_bytes(int64_t K, int64_t* bytes) {
  if (K < 0 || K > 0x7fffffffll || !bytes) return fail(A3D_ERR_ARG, "a3d_batch_order_ws_bytes: bad argument");
  const int64_t n = K > 0 ? K : 1;
  *bytes = 2 * al256(8 * n) + al256(4 * n) + al256((int64_t)sort_bytes(K));
  return A3D_OK;
}

extern "C" int a3d_batch_order(const void* const* blobs, int n_meshes, const float* P, int64_t K,
                               const int64_t* q_off, uint32_t* perm, void* ws, int64_t ws_bytes,
                               a3d_stream_t stream) {
  int64_t need = 0;
  if (int rc = a3d_batch_order_ws_bytes(K, &need)) return rc;
  if (n_meshes <= 0 || n_meshes > (1 << 24)) return fail(A3D_ERR_ARG, "a3d_batch_order: n_meshes out of range");
  if (K == 0) return A3D_OK;
  if (!blobs || !P || !q_off || !perm || !ws) return fail(A3D_ERR_ARG, "a3d_batch_order: null argument");
  if (ws_bytes < need) return fail(A3D_ERR_CAPACITY, "a3d_batch_order: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  char* w = (char*)ws;
  uint64_t* k0 = (uint64_t*)w;
  uint64_t* k1 = (uint64_t*)(w + al256(8 * K));
  uint32_t* v0 = (uint32_t*)(w + 2 * al256(8 * K));
  void* tmp = w + 2 * al256(8 * K) + al256(4 * K);
  size_t tb = sort_bytes(K);
  hipLaunchKernelGGL(k_batch_keys, dim3((unsigned)div_up(K, kBlkB)), dim3(kBlkB), 0, s, blobs, n_meshes, P, K, q_off,
                     k0, v0);
  A3D_LAUNCH_CHECK();
  int mbits = 0;
  while ((1ll << mbits) < n_meshes) ++mbits;
  if (mbits <= 8) {
    uint32_t* c0 = (uint32_t*)k1;  // the 64-bit key slots hold the 32-bit keys
    uint32_t* c1 = c0 + K;
    hipLaunchKernelGGL(k_batch_keys32, dim3((unsigned)div_up(K, kBlkB)), dim3(kBlkB), 0, s, k0, K, c0);
    A3D_LAUNCH_CHECK();
    A3D_HIP_CHECK(rocprim::radix_sort_pairs(tmp, tb, c0, c1, v0, perm, (int)K, 0, 24 + mbits, s));
    return A3D_OK;
  }
  A3D_HIP_CHECK(rocprim::radix_sort_pairs(tmp, tb, k0, k1, v0, perm, (int)K, 0, 30 + mbits, s));
  return A3D_OK;
}

extern "C" int a3d_batch_udf(const void* const* blobs, const float* P, const uint32_t* perm, const int64_t* tiles,
                             int64_t n_tiles, unsigned int* counter, int waves_per_block, int blocks, float* dist,
                             int32_t* face, float* closest, float* uvw, const void* bounds, a3d_stream_t stream) {
  if (n_tiles < 0 || n_tiles > 0xffffffffll) return fail(A3D_ERR_ARG, "a3d_batch_udf: n_tiles out of range");
  if (waves_per_block < 1 || waves_per_block > 16 || blocks < 1 || blocks > (1 << 20))
    return fail(A3D_ERR_ARG, "a3d_batch_udf: launch shape %d waves x %d blocks out of range", waves_per_block, blocks);
  if (n_tiles == 0) return A3D_OK;
  if (!blobs || !P || !perm || !tiles || !counter || !dist || !face || !closest)
    return fail(A3D_ERR_ARG, "a3d_batch_udf: null argument");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, counter);
  A3D_LAUNCH_CHECK();
  const size_t lds = (size_t)waves_per_block * (kWaveStack * 2 * sizeof(float4) + sizeof(WaveQ));
  const GridBound* gb = (const GridBound*)bounds;
  hipLaunchKernelGGL(k_udf_batch, dim3(blocks), dim3(64 * waves_per_block), lds, s, blobs, P, perm, tiles, n_tiles,
                     counter, dist, face, closest, uvw, gb, false);
  A3D_LAUNCH_CHECK();
  if (gb) {  // lanes whose bound fell short: walk them again without one
    hipLaunchKernelGGL(k_udf_batch, dim3(blocks), dim3(64 * waves_per_block), lds, s, blobs, P, perm, tiles, n_tiles,
                       counter, dist, face, closest, uvw, (const GridBound*)nullptr, true);
    A3D_LAUNCH_CHECK();
  }
  return A3D_OK;
}

