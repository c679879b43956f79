// Flood fill on a bit-packed voxel crop for gfx950.
//
// Replaces flood_fill_kernel + the host BFS loop (Atom3D
// atom3d/kernels/flood_fill_kernels.cu:21-175) and the dense post-processing
// of flood_fill_3d_sparse (kernels/flood_fill.py:118-258) and
// get_dam_neighborhood_signs (apps/sparse_flood_fill.py:34-100).
//
// The reference labels (-2 dry, -1 dam, 1 collision, 2 water) are a function
// of one set only: R, the free voxels connected to the start under 6/18/26
// connectivity.  dam = occupied & dilate(R); collision = R & dilate(occ);
// water = R \ collision; dry = the rest.  So instead of one launch, one device
// sync and one host read per BFS level (~1000 levels on a 1024^3 crop), R is
// computed as a monotone fixpoint on bit planes:
//   - planes hold 1 bit per voxel, 32 voxels along k per word, grouped in
//     bricks of 16 (i) x 16 (j) x 4 words (128 voxels along k) = 4 KB, stored
//     brick after brick so a brick loads as 256 coalesced 16-byte reads;
//   - one workgroup per active brick loads its words plus a one-word halo
//     into LDS and iterates R <- R | (dilate(R) & free) to a local fixpoint
//     (32 voxels per bitwise op, two barriers per step), each step also
//     filling every free k-run that holds a reached voxel (carry
//     propagation, krun_fill) so water crosses a brick along k in one step;
//   - a brick whose words changed re-activates its 26 neighbours for the next
//     global pass (flag = pass number, so nothing is cleared between passes);
//     the host checks for convergence once every 16 passes.
// Classification, the k-ring neighbourhood (26-dilations of the dam plane)
// and the lexicographic coordinate extraction (popcount, prefix sum, write)
// run on the same planes.
#include "a3d_prim.hpp"

#include "a3d_common.hpp"

namespace a3d {
namespace {

constexpr int kBlk = 256;
constexpr int kHI = 18, kHJ = 18, kHM = 6;  // halo brick: 16+2 x 16+2 x 4+2 words
constexpr int kHalo = kHI * kHJ * kHM;
constexpr int kMaxPasses = 1 << 16;
constexpr int kCheckEvery = 16;

enum Plane { P_OCC = 0, P_REACH, P_WATER, P_COL, P_DAM, P_DRY, P_EXT, P_NB, P_EXT_TIDE, P_DRY_ADJ, P_TMP, P_N };

struct Geo {
  int D, H, W;
  int Wn;             // words per row
  int nbi, nbj, nbk;  // bricks per axis
  int64_t words;      // words per plane
};

Geo make_geo(int D, int H, int W) {
  Geo g;
  g.D = D; g.H = H; g.W = W;
  g.Wn = (W + 31) / 32;
  g.nbi = (D + 15) / 16;
  g.nbj = (H + 15) / 16;
  g.nbk = (g.Wn + 3) / 4;
  g.words = (int64_t)g.nbi * g.nbj * g.nbk * 1024;
  return g;
}

__device__ __forceinline__ int64_t waddr(const Geo& g, int i, int j, int mw) {
  int64_t b = ((int64_t)(i >> 4) * g.nbj + (j >> 4)) * g.nbk + (mw >> 2);
  return (b << 10) + ((((i & 15) << 4) | (j & 15)) << 2) + (mw & 3);
}

__device__ __forceinline__ uint32_t valid_bits(const Geo& g, int i, int j, int mw) {
  if (i >= g.D || j >= g.H) return 0u;
  int k0 = mw * 32;
  if (k0 + 32 <= g.W) return 0xffffffffu;
  if (k0 >= g.W) return 0u;
  return (1u << (g.W - k0)) - 1u;
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  for (int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlk) p[t] = v;
}

__global__ void k_scatter(const int64_t* __restrict__ ijk, int64_t N, int64_t l0, int64_t l1, int64_t l2, Geo g,
                          uint32_t* occ) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= N) return;
  int64_t i = ijk[t * 3] - l0, j = ijk[t * 3 + 1] - l1, k = ijk[t * 3 + 2] - l2;
  if (i < 0 || j < 0 || k < 0 || i >= g.D || j >= g.H || k >= g.W) return;
  atomicOr(occ + waddr(g, (int)i, (int)j, (int)(k >> 5)), 1u << (k & 31));
}

// Dense bool occupancy [D,H,W] -> packed plane (one thread per plane word).
__global__ void k_pack_dense(const uint8_t* __restrict__ occ, Geo g, uint32_t* out) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  int64_t n = (int64_t)g.D * g.H * g.Wn;
  if (t >= n) return;
  int mw = (int)(t % g.Wn);
  int64_t r = t / g.Wn;
  int j = (int)(r % g.H), i = (int)(r / g.H);
  uint32_t w = 0;
  const uint8_t* row = occ + ((int64_t)i * g.H + j) * g.W;
  for (int b = 0; b < 32; ++b) {
    int k = mw * 32 + b;
    if (k < g.W && row[k]) w |= 1u << b;
  }
  out[waddr(g, i, j, mw)] = w;
}

__global__ void k_seed(Geo g, int si, int sj, int sk, const uint32_t* __restrict__ occ, uint32_t* reach,
                       int32_t* active, int32_t* flag) {
  if (threadIdx.x != 0) return;
  int64_t a = waddr(g, si, sj, sk >> 5);
  uint32_t bit = 1u << (sk & 31);
  if (occ[a] & bit) { *flag = 1; return; }
  *flag = 0;
  reach[a] |= bit;
  int64_t b = ((int64_t)(si >> 4) * g.nbj + (sj >> 4)) * g.nbk + (sk >> 7);
  active[b] = 0;
}

__device__ __forceinline__ void brick_of(const Geo& g, int64_t b, int& bi, int& bj, int& bk) {
  bk = (int)(b % g.nbk);
  int64_t r = b / g.nbk;
  bj = (int)(r % g.nbj);
  bi = (int)(r / g.nbj);
}

// Load a plane's brick plus its one-word halo into LDS: s[hi][hj][hm].
__device__ __forceinline__ void load_halo(const Geo& g, const uint32_t* __restrict__ plane, int bi, int bj, int bk,
                                          uint32_t* s) {
  for (int t = threadIdx.x; t < kHalo; t += kBlk) {
    int hm = t % kHM, r = t / kHM;
    int hj = r % kHJ, hi = r / kHJ;
    int i = bi * 16 + hi - 1, j = bj * 16 + hj - 1, mw = bk * 4 + hm - 1;
    uint32_t v = 0;
    if (i >= 0 && j >= 0 && mw >= 0 && i < g.nbi * 16 && j < g.nbj * 16 && mw < g.nbk * 4) v = plane[waddr(g, i, j, mw)];
    s[t] = v;
  }
}

__device__ __forceinline__ uint32_t hw(const uint32_t* s, int hi, int hj, int hm) { return s[(hi * kHJ + hj) * kHM + hm]; }

// Word m (0..3) of row (hi, hj) dilated by one voxel along k.
__device__ __forceinline__ uint32_t dk(const uint32_t* s, int hi, int hj, int m) {
  uint32_t w = hw(s, hi, hj, m + 1), l = hw(s, hi, hj, m), r = hw(s, hi, hj, m + 2);
  return w | (w << 1) | (l >> 31) | (w >> 1) | (r << 31);
}

// Neighbourhood dilation (including the voxel itself) of own word m of
// interior row (ii, jj): 6 = L1 <= 1, 18 = L1 <= 2, 26 = all.
__device__ __forceinline__ uint32_t dilate(const uint32_t* s, int ii, int jj, int m, int conn) {
  int hi = ii + 1, hj = jj + 1;
  uint32_t c = dk(s, hi, hj, m);
  if (conn == 6)
    return c | hw(s, hi - 1, hj, m + 1) | hw(s, hi + 1, hj, m + 1) | hw(s, hi, hj - 1, m + 1) | hw(s, hi, hj + 1, m + 1);
  uint32_t f = dk(s, hi - 1, hj, m) | dk(s, hi + 1, hj, m) | dk(s, hi, hj - 1, m) | dk(s, hi, hj + 1, m);
  if (conn == 18)
    return c | f | hw(s, hi - 1, hj - 1, m + 1) | hw(s, hi - 1, hj + 1, m + 1) | hw(s, hi + 1, hj - 1, m + 1) |
           hw(s, hi + 1, hj + 1, m + 1);
  return c | f | dk(s, hi - 1, hj - 1, m) | dk(s, hi - 1, hj + 1, m) | dk(s, hi + 1, hj - 1, m) |
         dk(s, hi + 1, hj + 1, m);
}

// Every free voxel of a 128-voxel k-row (4 words, bit b of word m = voxel
// 32 m + b) joined along k to a voxel of x: a run of free voxels holding a
// reached one is reached whole, whatever the connectivity (k steps are in all
// of them).  Up: adding the seeds to the free mask carries from the lowest
// seed of each run through the run's top, so ((F + S) ^ F) & F | S is each
// run from its lowest seed up; down: the same on the bit-reversed row.
__device__ __forceinline__ void krun_fill(uint32_t x[4], const uint32_t f[4]) {
  const uint64_t F0 = f[0] | ((uint64_t)f[1] << 32), F1 = f[2] | ((uint64_t)f[3] << 32);
  const uint64_t S0 = x[0] | ((uint64_t)x[1] << 32), S1 = x[2] | ((uint64_t)x[3] << 32);
  uint64_t a0 = F0 + S0;
  uint64_t a1 = F1 + S1 + (a0 < F0 ? 1ull : 0ull);
  const uint64_t U0 = ((a0 ^ F0) & F0) | S0, U1 = ((a1 ^ F1) & F1) | S1;
  const uint64_t rF0 = __builtin_bitreverse64(F1), rF1 = __builtin_bitreverse64(F0);
  const uint64_t rS0 = __builtin_bitreverse64(U1), rS1 = __builtin_bitreverse64(U0);
  a0 = rF0 + rS0;
  a1 = rF1 + rS1 + (a0 < rF0 ? 1ull : 0ull);
  const uint64_t D0 = ((a0 ^ rF0) & rF0) | rS0, D1 = ((a1 ^ rF1) & rF1) | rS1;
  const uint64_t X0 = __builtin_bitreverse64(D1), X1 = __builtin_bitreverse64(D0);
  x[0] = (uint32_t)X0;
  x[1] = (uint32_t)(X0 >> 32);
  x[2] = (uint32_t)X1;
  x[3] = (uint32_t)(X1 >> 32);
}

__global__ __launch_bounds__(kBlk) void k_ff_pass(Geo g, int conn, int t, const uint32_t* __restrict__ occ,
                                                   uint32_t* reach, int32_t* active, int32_t* changed) {
  const int64_t b = blockIdx.x;
  if (active[b] != t) return;
  __shared__ uint32_t s[kHalo];
  int bi, bj, bk;
  brick_of(g, b, bi, bj, bk);
  load_halo(g, reach, bi, bj, bk, s);
  const int ii = threadIdx.x >> 4, jj = threadIdx.x & 15;
  const int i = bi * 16 + ii, j = bj * 16 + jj;
  const int64_t base = (b << 10) + ((int64_t)threadIdx.x << 2);
  uint32_t fr[4], cur[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    fr[m] = ~occ[base + m] & valid_bits(g, i, j, bk * 4 + m);
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 4; ++m) cur[m] = hw(s, ii + 1, jj + 1, m + 1);
  bool any = false;
  while (true) {
    uint32_t nx[4];
    bool ch = false;
#pragma unroll
    for (int m = 0; m < 4; ++m) nx[m] = cur[m] | (dilate(s, ii, jj, m, conn) & fr[m]);
    krun_fill(nx, fr);  // along k a step reaches the whole free run, not one voxel
#pragma unroll
    for (int m = 0; m < 4; ++m) ch |= nx[m] != cur[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      cur[m] = nx[m];
      s[((ii + 1) * kHJ + (jj + 1)) * kHM + m + 1] = nx[m];
    }
    any |= ch;
    if (!__syncthreads_or(ch)) break;
  }
  if (__syncthreads_or(any)) {
#pragma unroll
    for (int m = 0; m < 4; ++m) reach[base + m] = cur[m];
    if (threadIdx.x == 0) atomicAdd(&changed[t], 1);
    if (threadIdx.x < 27) {
      int di = threadIdx.x / 9 - 1, dj = (threadIdx.x / 3) % 3 - 1, dd = threadIdx.x % 3 - 1;
      int ni = bi + di, nj = bj + dj, nk = bk + dd;
      if ((di | dj | dd) != 0 && ni >= 0 && nj >= 0 && nk >= 0 && ni < g.nbi && nj < g.nbj && nk < g.nbk)
        atomicMax(&active[((int64_t)ni * g.nbj + nj) * g.nbk + nk], t + 1);
    }
  }
}

// Labels from (occupancy, reached): the reference's mask, then the extended
// dam of flood_fill_3d_sparse (dry & occupied & dilated(flooded), dilation =
// 6-cross for connectivity 6, 3x3x3 otherwise).
__global__ __launch_bounds__(kBlk) void k_ff_classify(Geo g, int conn, uint32_t* planes) {
  const int64_t b = blockIdx.x;
  const int64_t W = g.words;
  __shared__ uint32_t sR[kHalo];
  __shared__ uint32_t sO[kHalo];
  int bi, bj, bk;
  brick_of(g, b, bi, bj, bk);
  load_halo(g, planes + P_REACH * W, bi, bj, bk, sR);
  load_halo(g, planes + P_OCC * W, bi, bj, bk, sO);
  __syncthreads();
  const int ii = threadIdx.x >> 4, jj = threadIdx.x & 15;
  const int i = bi * 16 + ii, j = bj * 16 + jj;
  const int64_t base = (b << 10) + ((int64_t)threadIdx.x << 2);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    uint32_t v = valid_bits(g, i, j, bk * 4 + m);
    uint32_t o = hw(sO, ii + 1, jj + 1, m + 1), r = hw(sR, ii + 1, jj + 1, m + 1);
    uint32_t dam = o & dilate(sR, ii, jj, m, conn);
    uint32_t col = r & dilate(sO, ii, jj, m, conn);
    uint32_t ext = o & ~dam & dilate(sR, ii, jj, m, conn == 6 ? 6 : 26);
    uint32_t damf = dam | ext;
    planes[P_WATER * W + base + m] = r & ~col;
    planes[P_COL * W + base + m] = col;
    planes[P_DAM * W + base + m] = damf;
    planes[P_EXT * W + base + m] = ext;
    planes[P_DRY * W + base + m] = v & ~r & ~damf;
  }
}

// dst = dilate26(src) restricted to the crop.
__global__ __launch_bounds__(kBlk) void k_ff_dilate26(Geo g, const uint32_t* __restrict__ src, uint32_t* dst) {
  const int64_t b = blockIdx.x;
  __shared__ uint32_t s[kHalo];
  int bi, bj, bk;
  brick_of(g, b, bi, bj, bk);
  load_halo(g, src, bi, bj, bk, s);
  __syncthreads();
  const int ii = threadIdx.x >> 4, jj = threadIdx.x & 15;
  const int64_t base = (b << 10) + ((int64_t)threadIdx.x << 2);
#pragma unroll
  for (int m = 0; m < 4; ++m)
    dst[base + m] = dilate(s, ii, jj, m, 26) & valid_bits(g, bi * 16 + ii, bj * 16 + jj, bk * 4 + m);
}

__global__ void k_ff_combine(int64_t W, uint32_t* planes) {
  for (int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x; t < W; t += (int64_t)gridDim.x * kBlk) {
    uint32_t nb = planes[P_NB * W + t];
    planes[P_EXT_TIDE * W + t] = nb & (planes[P_WATER * W + t] | planes[P_COL * W + t]);
    planes[P_DRY_ADJ * W + t] = nb & planes[P_DRY * W + t];
  }
}

__global__ void k_ff_totals(int64_t W, const uint32_t* __restrict__ planes, unsigned long long* tot) {
  int p = blockIdx.y;
  unsigned long long c = 0;
  for (int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x; t < W; t += (int64_t)gridDim.x * kBlk)
    c += __popc(planes[p * W + t]);
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(&tot[p], c);
}

// Row-major word counts for ordered extraction.
__global__ void k_ff_count(Geo g, const uint32_t* __restrict__ plane, uint32_t* __restrict__ cnt) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  int64_t n = (int64_t)g.D * g.H * g.Wn;
  if (t >= n) return;
  int mw = (int)(t % g.Wn);
  int64_t r = t / g.Wn;
  cnt[t] = __popc(plane[waddr(g, (int)(r / g.H), (int)(r % g.H), mw)]);
}

// One block per kBlk consecutive words; the block's set bits are written as
// consecutive output rows by consecutive threads (rank -> word by binary
// search over the block's prefix counts in LDS, then the bit by a select in
// the word), so the 24-byte rows are written coalesced.
__global__ __launch_bounds__(kBlk) void k_ff_write(Geo g, const uint32_t* __restrict__ plane,
                                                   const uint32_t* __restrict__ off, int64_t l0, int64_t l1,
                                                   int64_t l2, int64_t* __restrict__ out) {
  __shared__ uint32_t s_word[kBlk];
  __shared__ uint32_t s_pre[kBlk + 1];
  const int64_t n = (int64_t)g.D * g.H * g.Wn;
  const int64_t t0 = blockIdx.x * (int64_t)kBlk;
  const int64_t t = t0 + threadIdx.x;
  uint32_t w = 0;
  if (t < n) {
    const int mw = (int)(t % g.Wn);
    const int64_t r = t / g.Wn;
    w = plane[waddr(g, (int)(r / g.H), (int)(r % g.H), mw)];
  }
  const uint32_t base = off[t0];
  s_word[threadIdx.x] = w;
  s_pre[threadIdx.x] = t < n ? off[t] - base : 0;
  __syncthreads();
  const int64_t last = (n - t0 < kBlk ? n - t0 : kBlk) - 1;
  if (threadIdx.x == 0) s_pre[kBlk] = s_pre[last] + __popc(s_word[last]);
  __syncthreads();
  const uint32_t total = s_pre[kBlk];
  for (uint32_t e = threadIdx.x; e < total; e += kBlk) {
    int lo = 0, hi = (int)last;  // last word with s_pre <= e
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_pre[mid] <= e) lo = mid; else hi = mid - 1;
    }
    uint32_t x = s_word[lo];
    uint32_t k = e - s_pre[lo];  // the k-th set bit of x (0-based)
    int bit = 0;
    for (int width = 16; width > 0; width >>= 1) {
      const uint32_t m = x & ((1u << width) - 1u);
      const uint32_t c = __popc(m);
      if (k >= c) {
        k -= c;
        x >>= width;
        bit += width;
      } else {
        x = m;
      }
    }
    const int64_t tw = t0 + lo;
    const int mw = (int)(tw % g.Wn);
    const int64_t r = tw / g.Wn;
    const int64_t o = (int64_t)base + e;
    out[o * 3] = l0 + r / g.H;
    out[o * 3 + 1] = l1 + r % g.H;
    out[o * 3 + 2] = l2 + mw * 32 + bit;
  }
}

__global__ void k_ff_dense_mask(Geo g, const uint32_t* __restrict__ planes, int32_t* __restrict__ mask) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  int64_t n = (int64_t)g.D * g.H * g.W;
  if (t >= n) return;
  int k = (int)(t % g.W);
  int64_t r = t / g.W;
  int j = (int)(r % g.H), i = (int)(r / g.H);
  int64_t a = waddr(g, i, j, k >> 5);
  uint32_t bit = 1u << (k & 31);
  const int64_t W = g.words;
  int32_t v = -2;
  if (planes[P_WATER * W + a] & bit) v = 2;
  else if (planes[P_COL * W + a] & bit) v = 1;
  else if ((planes[P_DAM * W + a] & ~planes[P_EXT * W + a]) & bit) v = -1;
  mask[t] = v;
}

__global__ void k_ff_dam_mask(Geo g, const uint32_t* __restrict__ dam, const int64_t* __restrict__ ijk, int64_t N,
                              int64_t l0, int64_t l1, int64_t l2, uint8_t* __restrict__ out) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= N) return;
  int64_t i = ijk[t * 3] - l0, j = ijk[t * 3 + 1] - l1, k = ijk[t * 3 + 2] - l2;
  uint8_t v = 0;
  if (i >= 0 && j >= 0 && k >= 0 && i < g.D && j < g.H && k < g.W)
    v = (dam[waddr(g, (int)i, (int)j, (int)(k >> 5))] >> (k & 31)) & 1u;
  out[t] = v;
}

inline int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

struct FWs {
  int32_t* active;
  int32_t* changed;
  int32_t* flag;
  unsigned long long* totals;
  uint32_t* cnt;
  uint32_t* off;
  void* cub;
  size_t cub_bytes;
  int64_t total;
};

FWs carve_f(void* base, const Geo& g) {
  FWs w;
  char* b = (char*)base;
  int64_t o = 0;
  auto take = [&](int64_t bytes) { char* p = b ? b + o : nullptr; o += al256(bytes); return p; };
  int64_t nb = (int64_t)g.nbi * g.nbj * g.nbk;
  int64_t rows = (int64_t)g.D * g.H * g.Wn;
  int64_t nr = rows > 1 ? rows : 2;
  w.active = (int32_t*)take(4 * nb);
  w.changed = (int32_t*)take(4 * (int64_t)kMaxPasses);
  w.flag = (int32_t*)take(64);
  w.totals = (unsigned long long*)take(8 * P_N);
  w.cnt = (uint32_t*)take(4 * nr);
  w.off = (uint32_t*)take(4 * nr);
  size_t tb = 0;
  (void)a3d::exclusive_sum(nullptr, tb, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)nr);
  w.cub_bytes = tb;
  w.cub = take((int64_t)tb);
  w.total = o;
  return w;
}

int check_dims(int D, int H, int W, const char* who) {
  if (D <= 0 || H <= 0 || W <= 0) return fail(A3D_ERR_ARG, "%s: crop %dx%dx%d must be non-empty", who, D, H, W);
  if ((int64_t)D * H * W >= (int64_t(1) << 32))
    return fail(A3D_ERR_ARG, "%s: crop of %lld voxels exceeds 2^32", who, (long long)D * H * W);
  if ((int64_t)D * H * ((W + 31) / 32) >= (int64_t(1) << 31))
    return fail(A3D_ERR_ARG, "%s: crop too large", who);
  return A3D_OK;
}

inline unsigned gridn(int64_t n) { return (unsigned)div_up(n > 0 ? n : 1, kBlk); }

// Per-axis min and max of int64 [N,3] coordinates: a wave reduction, then
// one 64-bit atomic per wave and axis.
__global__ void k_ijk_bounds_init(long long* out) {
  if (threadIdx.x < 3) out[threadIdx.x] = 0x7fffffffffffffffll;
  else if (threadIdx.x < 6) out[threadIdx.x] = -0x7fffffffffffffffll - 1;
}

__global__ __launch_bounds__(256) void k_ijk_bounds(const int64_t* __restrict__ ijk, int64_t N,
                                                    long long* __restrict__ out) {
  long long lo[3] = {0x7fffffffffffffffll, 0x7fffffffffffffffll, 0x7fffffffffffffffll};
  long long hi[3] = {-0x7fffffffffffffffll - 1, -0x7fffffffffffffffll - 1, -0x7fffffffffffffffll - 1};
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < N; i += (int64_t)gridDim.x * 256) {
    for (int k = 0; k < 3; ++k) {
      long long v = ijk[i * 3 + k];
      lo[k] = v < lo[k] ? v : lo[k];
      hi[k] = v > hi[k] ? v : hi[k];
    }
  }
  for (int k = 0; k < 3; ++k) {
    for (int o = 32; o > 0; o >>= 1) {
      long long a = __shfl_xor(lo[k], o), b = __shfl_xor(hi[k], o);
      lo[k] = a < lo[k] ? a : lo[k];
      hi[k] = b > hi[k] ? b : hi[k];
    }
  }
  __shared__ long long s_lo[4][3], s_hi[4][3];
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 3; ++k) {
      s_lo[wv][k] = lo[k];
      s_hi[wv][k] = hi[k];
    }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    long long a = s_lo[0][k], b = s_hi[0][k];
    for (int w = 1; w < 4; ++w) {
      a = s_lo[w][k] < a ? s_lo[w][k] : a;
      b = s_hi[w][k] > b ? s_hi[w][k] : b;
    }
    atomicMin(&out[k], a);
    atomicMax(&out[3 + k], b);
  }
}

}  // namespace
}  // namespace a3d

using namespace a3d;

extern "C" int a3d_ijk_bounds(const int64_t* ijk, int64_t N, int64_t* out6, a3d_stream_t stream) {
  if (!out6 || (N > 0 && !ijk) || N < 0) return fail(A3D_ERR_ARG, "a3d_ijk_bounds: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_ijk_bounds_init, dim3(1), dim3(64), 0, s, (long long*)out6);
  A3D_LAUNCH_CHECK();
  if (N == 0) return A3D_OK;
  int64_t g = (N + 255) / 256;
  hipLaunchKernelGGL(k_ijk_bounds, dim3((unsigned)(g < 1024 ? g : 1024)), dim3(256), 0, s, ijk, N, (long long*)out6);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_flood_sizes(int D, int H, int W, int64_t* plane_words, int64_t* ws_bytes) {
  if (int rc = check_dims(D, H, W, "a3d_flood_sizes")) return rc;
  if (!plane_words || !ws_bytes) return fail(A3D_ERR_ARG, "a3d_flood_sizes: null output");
  Geo g = make_geo(D, H, W);
  *plane_words = g.words;
  *ws_bytes = carve_f(nullptr, g).total;
  return A3D_OK;
}

extern "C" int a3d_flood_pack_dense(const uint8_t* occ, int D, int H, int W, uint32_t* state, a3d_stream_t stream) {
  if (int rc = check_dims(D, H, W, "a3d_flood_pack_dense")) return rc;
  if (!occ || !state) return fail(A3D_ERR_ARG, "a3d_flood_pack_dense: null argument");
  Geo g = make_geo(D, H, W);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_u32, dim3(4096), dim3(kBlk), 0, s, state + P_OCC * g.words, g.words, 0u);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pack_dense, dim3(gridn((int64_t)D * H * g.Wn)), dim3(kBlk), 0, s, occ, g, state + P_OCC * g.words);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_flood_fill(const int64_t* dam_ijk, int64_t N, const int64_t* lo_host, int D, int H, int W,
                              int si, int sj, int sk, int conn, int k_ring, uint32_t* state, void* ws,
                              int64_t ws_bytes, int64_t* totals_host, int64_t* passes_host, a3d_stream_t stream) {
  if (int rc = check_dims(D, H, W, "a3d_flood_fill")) return rc;
  if (conn != 6 && conn != 18 && conn != 26) return fail(A3D_ERR_ARG, "a3d_flood_fill: connectivity %d not in {6,18,26}", conn);
  if (k_ring < 0 || k_ring > 64) return fail(A3D_ERR_ARG, "a3d_flood_fill: k_ring %d not in [0, 64]", k_ring);
  if (!state || !ws || !totals_host || !passes_host) return fail(A3D_ERR_ARG, "a3d_flood_fill: null argument");
  if (dam_ijk && (N < 0 || !lo_host)) return fail(A3D_ERR_ARG, "a3d_flood_fill: bad dam list");
  if (si < 0 || sj < 0 || sk < 0 || si >= D || sj >= H || sk >= W)
    return fail(A3D_ERR_ARG, "a3d_flood_fill: start (%d,%d,%d) outside the %dx%dx%d crop", si, sj, sk, D, H, W);
  Geo g = make_geo(D, H, W);
  FWs w = carve_f(ws, g);
  if (ws_bytes < w.total) return fail(A3D_ERR_CAPACITY, "a3d_flood_fill: workspace %lld < %lld", (long long)ws_bytes, (long long)w.total);
  hipStream_t s = (hipStream_t)stream;
  const int64_t PW = g.words;
  const int64_t nb = (int64_t)g.nbi * g.nbj * g.nbk;
  // clear every plane but the occupancy when it was packed by the caller
  int64_t first = dam_ijk ? 0 : 1;
  hipLaunchKernelGGL(k_fill_u32, dim3(8192), dim3(kBlk), 0, s, state + first * PW, (P_N - first) * PW, 0u);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(kBlk), 0, s, (uint32_t*)w.active, nb, 0xffffffffu);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(kBlk), 0, s, (uint32_t*)w.changed, (int64_t)kMaxPasses, 0u);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlk), 0, s, (uint32_t*)w.totals, (int64_t)2 * P_N, 0u);
  A3D_LAUNCH_CHECK();
  if (dam_ijk && N > 0) {
    hipLaunchKernelGGL(k_scatter, dim3(gridn(N)), dim3(kBlk), 0, s, dam_ijk, N, lo_host[0], lo_host[1], lo_host[2], g,
                       state + P_OCC * PW);
    A3D_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_seed, dim3(1), dim3(64), 0, s, g, si, sj, sk, state + P_OCC * PW, state + P_REACH * PW,
                     w.active, w.flag);
  A3D_LAUNCH_CHECK();
  int32_t occupied = 0;
  A3D_HIP_CHECK(hipMemcpyAsync(&occupied, w.flag, 4, hipMemcpyDeviceToHost, s));
  A3D_HIP_CHECK(hipStreamSynchronize(s));
  if (occupied) {
    *passes_host = -1;  // start voxel is occupied: the reference returns empty results
    for (int p = 0; p < P_N; ++p) totals_host[p] = 0;
    return A3D_OK;
  }
  int t = 0;
  while (true) {
    if (t + kCheckEvery >= kMaxPasses) return fail(A3D_ERR_CAPACITY, "a3d_flood_fill: no convergence in %d passes", kMaxPasses);
    for (int q = 0; q < kCheckEvery; ++q, ++t) {
      hipLaunchKernelGGL(k_ff_pass, dim3((unsigned)nb), dim3(kBlk), 0, s, g, conn, t, state + P_OCC * PW,
                         state + P_REACH * PW, w.active, w.changed);
      A3D_LAUNCH_CHECK();
    }
    int32_t last = 0;
    A3D_HIP_CHECK(hipMemcpyAsync(&last, w.changed + (t - 1), 4, hipMemcpyDeviceToHost, s));
    A3D_HIP_CHECK(hipStreamSynchronize(s));
    if (last == 0) break;
  }
  *passes_host = t;
  hipLaunchKernelGGL(k_ff_classify, dim3((unsigned)nb), dim3(kBlk), 0, s, g, conn, state);
  A3D_LAUNCH_CHECK();
  // k-ring neighbourhood of the (final) dam: k 26-dilations
  const uint32_t* src = state + P_DAM * PW;
  for (int r = 0; r < k_ring; ++r) {
    uint32_t* dst = state + ((k_ring - r) % 2 == 1 ? P_NB : P_TMP) * PW;
    hipLaunchKernelGGL(k_ff_dilate26, dim3((unsigned)nb), dim3(kBlk), 0, s, g, src, dst);
    A3D_LAUNCH_CHECK();
    src = dst;
  }
  if (k_ring == 0) {
    hipLaunchKernelGGL(k_fill_u32, dim3(4096), dim3(kBlk), 0, s, state + P_NB * PW, PW, 0u);
    A3D_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_ff_combine, dim3(8192), dim3(kBlk), 0, s, PW, state);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ff_totals, dim3(1024, P_N), dim3(kBlk), 0, s, PW, state, w.totals);
  A3D_LAUNCH_CHECK();
  unsigned long long tot[P_N];
  A3D_HIP_CHECK(hipMemcpyAsync(tot, w.totals, sizeof(tot), hipMemcpyDeviceToHost, s));
  A3D_HIP_CHECK(hipStreamSynchronize(s));
  for (int p = 0; p < P_N; ++p) totals_host[p] = (int64_t)tot[p];
  return A3D_OK;
}

extern "C" int a3d_flood_extract(const uint32_t* state, int D, int H, int W, int plane, const int64_t* lo_host,
                                 int64_t* out_ijk, void* ws, int64_t ws_bytes, a3d_stream_t stream) {
  if (int rc = check_dims(D, H, W, "a3d_flood_extract")) return rc;
  if (plane < 0 || plane >= P_N) return fail(A3D_ERR_ARG, "a3d_flood_extract: plane %d out of range", plane);
  if (!state || !lo_host || !out_ijk || !ws) return fail(A3D_ERR_ARG, "a3d_flood_extract: null argument");
  Geo g = make_geo(D, H, W);
  FWs w = carve_f(ws, g);
  if (ws_bytes < w.total) return fail(A3D_ERR_CAPACITY, "a3d_flood_extract: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  int64_t rows = (int64_t)D * H * g.Wn;
  const uint32_t* pl = state + (int64_t)plane * g.words;
  hipLaunchKernelGGL(k_ff_count, dim3(gridn(rows)), dim3(kBlk), 0, s, g, pl, w.cnt);
  A3D_LAUNCH_CHECK();
  size_t tb = w.cub_bytes;
  A3D_HIP_CHECK(a3d::exclusive_sum(w.cub, tb, w.cnt, w.off, (int)rows, s));
  hipLaunchKernelGGL(k_ff_write, dim3(gridn(rows)), dim3(kBlk), 0, s, g, pl, w.off, lo_host[0], lo_host[1], lo_host[2],
                     out_ijk);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_flood_dense_mask(const uint32_t* state, int D, int H, int W, int32_t* mask, a3d_stream_t stream) {
  if (int rc = check_dims(D, H, W, "a3d_flood_dense_mask")) return rc;
  if (!state || !mask) return fail(A3D_ERR_ARG, "a3d_flood_dense_mask: null argument");
  Geo g = make_geo(D, H, W);
  hipLaunchKernelGGL(k_ff_dense_mask, dim3(gridn((int64_t)D * H * W)), dim3(kBlk), 0, (hipStream_t)stream, g, state, mask);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_flood_dam_mask(const uint32_t* state, int D, int H, int W, const int64_t* dam_ijk, int64_t N,
                                  const int64_t* lo_host, uint8_t* out, a3d_stream_t stream) {
  if (int rc = check_dims(D, H, W, "a3d_flood_dam_mask")) return rc;
  if (N < 0) return fail(A3D_ERR_ARG, "a3d_flood_dam_mask: N < 0");
  if (N == 0) return A3D_OK;
  if (!state || !dam_ijk || !lo_host || !out) return fail(A3D_ERR_ARG, "a3d_flood_dam_mask: null argument");
  Geo g = make_geo(D, H, W);
  hipLaunchKernelGGL(k_ff_dam_mask, dim3(gridn(N)), dim3(kBlk), 0, (hipStream_t)stream, g, state + P_DAM * g.words,
                     dam_ijk, N, lo_host[0], lo_host[1], lo_host[2], out);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}
