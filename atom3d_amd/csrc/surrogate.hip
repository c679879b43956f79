// Learned UDF surrogate for gfx950: a multiresolution hash grid (16 levels x
// 2^19 entries x 2 features) feeding an MLP 32 -> 64 -> 64 -> 1 (ReLU), with
// the matrix products on MFMA (v_mfma_f32_32x32x8_bf16: bf16 operands, fp32
// accumulation), trained on the device by the hand-written forward/backward
// below, and a router that answers far-field queries from the network and
// sends every query predicted inside a safety band to the exact BVH walk.
// There is no counterpart in the reference (it has no learned path).
//
// MFMA 32x32x8 lane layout (lane l, g = l / 32, n = l % 32, v = 0..15):
//   A[m][k]: lane l holds A[n][4g + j], j = 0..3        (bf16x4)
//   B[k][n]: lane l holds B[4g + j][n]                    (bf16x4)
//   D[m][n]: lane l holds D[8 (v / 4) + 4g + v % 4][n]    (f32 x16)
// Every layer is computed transposed, Y^T = W X^T, with the 32 queries of a
// tile as the N dimension: then the D registers of one layer are, after ReLU
// and rounding, exactly the B operand of the next (K-step 4 mb + v / 4,
// element v % 4), so activations never leave registers in the forward pass.
// Lane l computes the hash-grid features of query n for levels 4s + 2g and
// 4s + 2g + 1 (s = 0..3), which is the layer-1 B operand, and receives in the
// backward pass (dX^T = W1^T dH1) the gradient of exactly those features.
#include <hip/hip_fp16.h>
#include "a3d_prim.hpp"

#include <cmath>
#include <cstring>

#include "a3d_common.hpp"

namespace a3d {
namespace {

constexpr int kL = 16;
constexpr int kLog2T = 19;
constexpr int kT = 1 << kLog2T;
constexpr int kIn = 32;
constexpr int kH = 64;
constexpr int kBlkS = 256;
constexpr int kWavesS = kBlkS / 64;
// MLP parameter vector (fp32 master): W1 [64][32], W2 [64][64], w3 [64], b3,
// then the loss slot in gradient / partial vectors.
constexpr int kOffW1 = 0;
constexpr int kOffW2 = kOffW1 + kH * kIn;
constexpr int kOffW3 = kOffW2 + kH * kH;
constexpr int kOffB3 = kOffW3 + kH;
constexpr int kNMlp = kOffB3 + 1;  // 6209 trained parameters
constexpr int kOffLoss = kNMlp;
constexpr int kStride = 6272;      // partial / gradient vector stride (floats)
constexpr int kMaxTrainWaves = 1024;
constexpr uint32_t kSgMagic = 0x41335347;  // 'A3SG'

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct SgMeta {
  int32_t magic, levels, log2_t, features;
  float lo[3], inv_ext[3];
  float scale[kL];   // grid resolution N_l of each level
  int32_t res1[kL];  // N_l + 1 (dense stride)
  int32_t dense[kL]; // (N_l + 1)^3 <= T: direct indexing, else hashed
  float b3;
  int32_t pad[69];
};
static_assert(sizeof(SgMeta) == 512, "meta is 512 bytes");

constexpr int64_t kTabEntries = (int64_t)kL * kT;  // 2 features each
constexpr int64_t kOffTab16 = 512;
constexpr int64_t kOffW1b = kOffTab16 + kTabEntries * 2 * 2;
constexpr int64_t kOffW2b = kOffW1b + kH * kIn * 2;
constexpr int64_t kOffW3f = kOffW2b + kH * kH * 2;
constexpr int64_t kParamsBytes = kOffW3f + kH * 4;

struct SgState {
  float* tab;   // [L*T*2] fp32 master
  float* tm;    // Adam moments
  float* tv;
  float* tg;    // gradient accumulator
  float* mlp;   // [kStride] fp32 master
  float* mm;
  float* mv;
  float* mg;    // reduced gradient (+ loss at kOffLoss)
  float* part;  // [kMaxTrainWaves][kStride]
  int64_t bytes;
};

inline int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

SgState carve_s(void* base) {
  SgState s;
  char* b = (char*)base;
  int64_t o = 0;
  auto take = [&](int64_t n) { char* p = b ? b + o : nullptr; o += al256(n); return p; };
  const int64_t tb = kTabEntries * 2 * 4;
  s.tab = (float*)take(tb);
  s.tm = (float*)take(tb);
  s.tv = (float*)take(tb);
  s.tg = (float*)take(tb);
  s.mlp = (float*)take(kStride * 4);
  s.mm = (float*)take(kStride * 4);
  s.mv = (float*)take(kStride * 4);
  s.mg = (float*)take(kStride * 4);
  s.part = (float*)take((int64_t)kMaxTrainWaves * kStride * 4);
  s.bytes = o;
  return s;
}

__device__ __forceinline__ const SgMeta* meta_of(const void* params) { return (const SgMeta*)params; }
__device__ __forceinline__ const __half2* tab16_of(const void* params) {
  return (const __half2*)((const char*)params + kOffTab16);
}

// round to nearest even (finite values)
__device__ __forceinline__ short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

__device__ __forceinline__ uint32_t hash_idx(int ix, int iy, int iz) {
  return ((uint32_t)ix ^ ((uint32_t)iy * 2654435761u) ^ ((uint32_t)iz * 805459861u)) & (uint32_t)(kT - 1);
}

// Corner c (bit 0 x, 1 y, 2 z) of the cell containing x at level lv.
struct Cell {
  int i0[3];
  float w[3];
};

__device__ __forceinline__ Cell cell_of(const SgMeta& M, int lv, const float x[3]) {
  Cell c;
  const float N = M.scale[lv];
  for (int d = 0; d < 3; ++d) {
    float pos = x[d] * N;
    int i = (int)pos;
    const int top = (int)N - 1;
    i = i < 0 ? 0 : (i > top ? top : i);
    c.i0[d] = i;
    c.w[d] = pos - (float)i;
  }
  return c;
}

__device__ __forceinline__ uint32_t corner_idx(const SgMeta& M, int lv, const Cell& c, int k) {
  const int ix = c.i0[0] + (k & 1), iy = c.i0[1] + ((k >> 1) & 1), iz = c.i0[2] + ((k >> 2) & 1);
  if (M.dense[lv]) return (uint32_t)(ix + M.res1[lv] * (iy + M.res1[lv] * iz));
  return hash_idx(ix, iy, iz);
}

__device__ __forceinline__ float corner_w(const Cell& c, int k) {
  const float wx = (k & 1) ? c.w[0] : 1.0f - c.w[0];
  const float wy = (k & 2) ? c.w[1] : 1.0f - c.w[1];
  const float wz = (k & 4) ? c.w[2] : 1.0f - c.w[2];
  return (wx * wy) * wz;
}

__device__ __forceinline__ void unit_pos(const SgMeta& M, const float p[3], float x[3]) {
  for (int d = 0; d < 3; ++d) {
    float t = (p[d] - M.lo[d]) * M.inv_ext[d];
    x[d] = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  }
}

// The lane's 16 encoding features as the 4 layer-1 B operands.
template <bool F32TAB>
__device__ __forceinline__ void encode(const SgMeta& M, const void* tab, const float x[3], int g, bf16x4 xb[4]) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int lv = 4 * s + 2 * g + e;
      const Cell c = cell_of(M, lv, x);
      float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint32_t id = corner_idx(M, lv, c, k);
        const float w = corner_w(c, k);
        float2 v;
        if (F32TAB) {
          v = ((const float2*)tab)[(int64_t)lv * kT + id];
        } else {
          v = __half22float2(((const __half2*)tab)[(int64_t)lv * kT + id]);
        }
        f0 += w * v.x;
        f1 += w * v.y;
      }
      xb[s][2 * e] = f2bf(f0);
      xb[s][2 * e + 1] = f2bf(f1);
    }
    asm volatile("" ::: "memory");  // at most two levels (16 gathers) in flight
  }
}

__device__ __forceinline__ f32x16 mfma(bf16x4 a, bf16x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x16 zero16() {
  f32x16 z;
#pragma unroll
  for (int v = 0; v < 16; ++v) z[v] = 0.0f;
  return z;
}

// Weights in LDS (bf16) for the whole block.
struct WLds {
  short w1[kH * kIn];
  short w2[kH * kH];
  float w3[kH];
};

__device__ __forceinline__ void load_weights(const void* params, WLds& W) {
  const short* w1 = (const short*)((const char*)params + kOffW1b);
  const short* w2 = (const short*)((const char*)params + kOffW2b);
  const float* w3 = (const float*)((const char*)params + kOffW3f);
  for (int i = threadIdx.x; i < kH * kIn; i += blockDim.x) W.w1[i] = w1[i];
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) W.w2[i] = w2[i];
  for (int i = threadIdx.x; i < kH; i += blockDim.x) W.w3[i] = w3[i];
}

// Forward of one tile: h1, h2 pre-activations (D layout) and the layer-2
// B operands (post-ReLU bf16 h1).
__device__ __forceinline__ void mlp_forward(const WLds& W, int n, int g, const bf16x4 xb[4], f32x16 h1[2],
                                            bf16x4 hb[8], f32x16 h2[2]) {
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) {
    h1[mb] = zero16();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      bf16x4 a;
#pragma unroll
      for (int j = 0; j < 4; ++j) a[j] = W.w1[(32 * mb + n) * kIn + 8 * s + 4 * g + j];
      h1[mb] = mfma(a, xb[s], h1[mb]);
    }
  }
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int v = 0; v < 16; ++v) hb[4 * mb + v / 4][v % 4] = f2bf(fmaxf(h1[mb][v], 0.0f));
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) {
    h2[mb] = zero16();
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      bf16x4 a;
#pragma unroll
      for (int j = 0; j < 4; ++j) a[j] = W.w2[(32 * mb + n) * kH + 8 * kk + 4 * g + j];
      h2[mb] = mfma(a, hb[kk], h2[mb]);
    }
  }
}

__device__ __forceinline__ int hid(int mb, int v, int g) { return 32 * mb + 8 * (v / 4) + 4 * g + (v % 4); }

// Inference forward with one accumulator live at a time: layer-1 blocks turn
// into layer-2 operands as they finish, layer-2 blocks fold into the output.
__device__ __forceinline__ float mlp_infer(const WLds& W, int n, int g, const bf16x4 xb[4], float b3) {
  bf16x4 hb[8];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) {
    f32x16 acc = zero16();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      bf16x4 a;
#pragma unroll
      for (int j = 0; j < 4; ++j) a[j] = W.w1[(32 * mb + n) * kIn + 8 * s + 4 * g + j];
      acc = mfma(a, xb[s], acc);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) hb[4 * mb + v / 4][v % 4] = f2bf(fmaxf(acc[v], 0.0f));
  }
  float y = 0.0f;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) {
    f32x16 acc = zero16();
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      bf16x4 a;
#pragma unroll
      for (int j = 0; j < 4; ++j) a[j] = W.w2[(32 * mb + n) * kH + 8 * kk + 4 * g + j];
      acc = mfma(a, hb[kk], acc);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) y += W.w3[hid(mb, v, g)] * fmaxf(acc[v], 0.0f);
  }
  y += __shfl_xor(y, 32);
  return y + b3;
}

__device__ __forceinline__ float mlp_out(const WLds& W, int g, const f32x16 h2[2], float b3) {
  float y = 0.0f;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int v = 0; v < 16; ++v) y += W.w3[hid(mb, v, g)] * fmaxf(h2[mb][v], 0.0f);
  y += __shfl_xor(y, 32);
  return y + b3;
}

// ------------------------------------------------------------------ inference
// out[q] = network(P[q]) for the queries perm[i] (or i), one wave per tile of
// 32.  far (optional): for predictions >= band also writes face -1, closest =
// the query point and uvw 0, the outputs of a surrogate-answered query.
__global__ __launch_bounds__(kBlkS) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_sg_infer(const void* __restrict__ params, const float* __restrict__ P,
                                                    const uint32_t* __restrict__ perm, int64_t K,
                                                    float* __restrict__ out, float band, uint8_t* __restrict__ near,
                                                    int32_t* __restrict__ face, float* __restrict__ closest,
                                                    float* __restrict__ uvw) {
  __shared__ WLds W;
  load_weights(params, W);
  __syncthreads();
  const SgMeta& M = *meta_of(params);
  const __half2* tab = tab16_of(params);
  const int lane = threadIdx.x & 63, n = lane & 31, g = lane >> 5;
  const int64_t ntiles = (K + 31) / 32;
  for (int64_t tile = (int64_t)blockIdx.x * kWavesS + (threadIdx.x >> 6); tile < ntiles;
       tile += (int64_t)gridDim.x * kWavesS) {
    const int64_t i = tile * 32 + n;
    const bool valid = i < K;
    const int64_t q = valid ? (perm ? (int64_t)perm[i] : i) : 0;
    float p[3] = {0.0f, 0.0f, 0.0f};
    if (valid)
      for (int d = 0; d < 3; ++d) p[d] = P[q * 3 + d];
    float x[3];
    unit_pos(M, p, x);
    bf16x4 xb[4];
    encode<false>(M, tab, x, g, xb);
    asm volatile("" ::: "memory");  // keep the weight fragments in LDS, not in registers
    const float y = mlp_infer(W, n, g, xb, M.b3);
    if (valid && g == 0) {
      out[q] = y;
      if (near) near[i] = y < band ? 1 : 0;
      if (face && !(y < band)) {
        face[q] = -1;
        for (int d = 0; d < 3; ++d) {
          closest[q * 3 + d] = p[d];
          if (uvw) uvw[q * 3 + d] = 0.0f;
        }
      }
    }
  }
}

// ------------------------------------------------------------------ training
// One wave per tile of 32 samples (grid-stride): forward from the fp32 master
// table, loss (y - t)^2 / B, backward through the MLP on MFMA, hash-table
// gradients scattered with atomics, MLP weight gradients accumulated in
// registers over the wave's tiles (dW = dH X summed over the 32 queries of a
// tile is one more MFMA chain, its operands transposed through LDS), one
// partial gradient vector written per wave.
struct TLds {
  short dh2[kH * 32];  // [h][q]
  short h1[kH * 32];
  short dh1[kH * 32];
  short x[kIn * 32];   // [f][q]
  float red[kH * 32];  // dW3 products [h][q]
};

__device__ __forceinline__ bf16x4 lds4(const short* p) {
  bf16x4 r;
#pragma unroll
  for (int j = 0; j < 4; ++j) r[j] = p[j];
  return r;
}

__global__ __launch_bounds__(kBlkS) void k_sg_train(const void* __restrict__ params, const float* __restrict__ tab32,
                                                    const float* __restrict__ P, const float* __restrict__ target,
                                                    int64_t B, float inv_b, float* __restrict__ tg,
                                                    float* __restrict__ part) {
  __shared__ WLds W;
  __shared__ TLds T[kWavesS];
  load_weights(params, W);
  __syncthreads();
  const SgMeta& M = *meta_of(params);
  const int lane = threadIdx.x & 63, n = lane & 31, g = lane >> 5, wv = threadIdx.x >> 6;
  TLds& L = T[wv];
  const int64_t wave_id = (int64_t)blockIdx.x * kWavesS + wv;
  const int64_t nwaves = (int64_t)gridDim.x * kWavesS;
  const int64_t ntiles = (B + 31) / 32;
  f32x16 aw1[2], aw2[2][2];
  aw1[0] = zero16();
  aw1[1] = zero16();
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) aw2[a][b] = zero16();
  for (int i = lane; i < kH * 32; i += 64) L.red[i] = 0.0f;
  float ab3 = 0.0f, aloss = 0.0f;

  for (int64_t tile = wave_id; tile < ntiles; tile += nwaves) {
    const int64_t i = tile * 32 + n;
    const bool valid = i < B;
    float p[3] = {0.0f, 0.0f, 0.0f};
    float t = 0.0f;
    if (valid) {
      for (int d = 0; d < 3; ++d) p[d] = P[i * 3 + d];
      t = target[i];
    }
    float x[3];
    unit_pos(M, p, x);
    bf16x4 xb[4];
    encode<true>(M, tab32, x, g, xb);
    f32x16 h1[2], h2[2];
    bf16x4 hb[8];
    mlp_forward(W, n, g, xb, h1, hb, h2);
    const float y = mlp_out(W, g, h2, M.b3);
    const float diff = valid ? y - t : 0.0f;
    const float dy = 2.0f * diff * inv_b;
    if (g == 0) {
      aloss += diff * diff * inv_b;
      ab3 += dy;
    }
    // dH2 = dy w3 relu'(h2); dW3 += dy relu(h2)
    bf16x4 db[8];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int h = hid(mb, v, g);
        const bool on = h2[mb][v] > 0.0f;
        L.red[h * 32 + n] += on ? dy * h2[mb][v] : 0.0f;
        db[4 * mb + v / 4][v % 4] = f2bf(on ? dy * W.w3[h] : 0.0f);
      }
    // dH1 = W2^T dH2, masked by relu'(h1)
    f32x16 d1[2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      d1[mb] = zero16();
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        bf16x4 a;
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] = W.w2[(8 * s + 4 * g + j) * kH + 32 * mb + n];
        d1[mb] = mfma(a, db[s], d1[mb]);
      }
    }
    bf16x4 d1b[8];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int v = 0; v < 16; ++v) d1b[4 * mb + v / 4][v % 4] = f2bf(h1[mb][v] > 0.0f ? d1[mb][v] : 0.0f);
    // dX^T = W1^T dH1: the lane's own 16 features
    f32x16 dx = zero16();
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      bf16x4 a;
#pragma unroll
      for (int j = 0; j < 4; ++j) a[j] = W.w1[(8 * s + 4 * g + j) * kIn + n];
      dx = mfma(a, d1b[s], dx);
    }
    if (valid) {
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int lv = 4 * s + 2 * g + e;
          const float g0 = dx[4 * s + 2 * e], g1 = dx[4 * s + 2 * e + 1];
          if (g0 == 0.0f && g1 == 0.0f) continue;
          const Cell c = cell_of(M, lv, x);
          for (int k = 0; k < 8; ++k) {
            const uint32_t id = corner_idx(M, lv, c, k);
            const float w = corner_w(c, k);
            float* dst = tg + ((int64_t)lv * kT + id) * 2;
            atomicAdd(dst, w * g0);
            atomicAdd(dst + 1, w * g1);
          }
        }
    }
    // weight gradients: transpose dH2, relu(h1), dH1, X through LDS
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int h = hid(mb, v, g);
        L.dh2[h * 32 + n] = db[4 * mb + v / 4][v % 4];
        L.h1[h * 32 + n] = hb[4 * mb + v / 4][v % 4];
        L.dh1[h * 32 + n] = d1b[4 * mb + v / 4][v % 4];
      }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int j = 0; j < 4; ++j) L.x[(8 * s + 4 * g + j) * 32 + n] = xb[s][j];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // dW2[o][i] += sum_q dH2[o][q] h1[i][q]; dW1[h][f] += sum_q dH1[h][q] x[f][q]
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int q0 = 8 * s + 4 * g;
      bf16x4 xq = lds4(&L.x[n * 32 + q0]);
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        bf16x4 ad2 = lds4(&L.dh2[(32 * mb + n) * 32 + q0]);
        bf16x4 ad1 = lds4(&L.dh1[(32 * mb + n) * 32 + q0]);
        aw1[mb] = mfma(ad1, xq, aw1[mb]);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) aw2[mb][nb] = mfma(ad2, lds4(&L.h1[(32 * nb + n) * 32 + q0]), aw2[mb][nb]);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // write this wave's partial gradient vector
  float* out = part + wave_id * kStride;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int h = hid(mb, v, g);
      out[kOffW1 + h * kIn + n] = aw1[mb][v];
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) out[kOffW2 + h * kH + 32 * nb + n] = aw2[mb][nb][v];
    }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  {
    float s = 0.0f;
    for (int q = 0; q < 32; ++q) s += L.red[lane * 32 + q];
    out[kOffW3 + lane] = s;
  }
  for (int o = 16; o > 0; o >>= 1) {
    ab3 += __shfl_xor(ab3, o);
    aloss += __shfl_xor(aloss, o);
  }
  if (lane == 0) {
    out[kOffB3] = ab3;
    out[kOffLoss] = aloss;
  }
}

// g[p] = sum over the waves' partials (fixed order: 16 interleaved partial
// sums per parameter, then their sum).
__global__ __launch_bounds__(1024) void k_sg_reduce(const float* __restrict__ part, int64_t nparts,
                                                    float* __restrict__ mg) {
  __shared__ float red[16][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int p = blockIdx.x * 64 + tx;
  float s = 0.0f;
  if (p <= kOffLoss)
    for (int64_t w = ty; w < nparts; w += 16) s += part[w * kStride + p];
  red[ty][tx] = s;
  __syncthreads();
  if (ty == 0 && p <= kOffLoss) {
    float t = 0.0f;
    for (int k = 0; k < 16; ++k) t += red[k][tx];
    mg[p] = t;
  }
}

struct AdamArgs {
  float lr_t;  // lr * sqrt(1 - b2^t) / (1 - b1^t)
  float b1, b2, eps;
};

__device__ __forceinline__ float adam(float& m, float& v, float gr, const AdamArgs& a) {
  m = a.b1 * m + (1.0f - a.b1) * gr;
  v = a.b2 * v + (1.0f - a.b2) * gr * gr;
  return a.lr_t * m / (sqrtf(v) + a.eps);
}

// Hash table: entries with a zero gradient are left untouched (sparse Adam).
__global__ void k_sg_adam_table(void* params, float* __restrict__ tab, float* __restrict__ tm, float* __restrict__ tv,
                                float* __restrict__ tg, AdamArgs a) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= kTabEntries) return;
  const float g0 = tg[2 * e], g1 = tg[2 * e + 1];
  if (g0 == 0.0f && g1 == 0.0f) return;
  float m0 = tm[2 * e], m1 = tm[2 * e + 1], v0 = tv[2 * e], v1 = tv[2 * e + 1];
  const float p0 = tab[2 * e] - adam(m0, v0, g0, a);
  const float p1 = tab[2 * e + 1] - adam(m1, v1, g1, a);
  tab[2 * e] = p0;
  tab[2 * e + 1] = p1;
  tm[2 * e] = m0;
  tm[2 * e + 1] = m1;
  tv[2 * e] = v0;
  tv[2 * e + 1] = v1;
  tg[2 * e] = 0.0f;
  tg[2 * e + 1] = 0.0f;
  ((__half2*)((char*)params + kOffTab16))[e] = __floats2half2_rn(p0, p1);
}

// MLP: Adam on the reduced gradient, then the bf16 / fp32 copies the kernels read.
__global__ void k_sg_adam_mlp(void* params, float* __restrict__ mlp, float* __restrict__ mm, float* __restrict__ mv,
                              const float* __restrict__ mg, AdamArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNMlp) return;
  float m = mm[p], v = mv[p];
  const float w = mlp[p] - adam(m, v, mg[p], a);
  mlp[p] = w;
  mm[p] = m;
  mv[p] = v;
  char* pb = (char*)params;
  if (p < kOffW2) ((short*)(pb + kOffW1b))[p - kOffW1] = f2bf(w);
  else if (p < kOffW3) ((short*)(pb + kOffW2b))[p - kOffW2] = f2bf(w);
  else if (p < kOffB3) ((float*)(pb + kOffW3f))[p - kOffW3] = w;
  else ((SgMeta*)pb)->b3 = w;
}

__device__ __forceinline__ uint32_t pcg(uint64_t& s) {
  s = s * 6364136223846793005ull + 1442695040888963407ull;
  uint32_t x = (uint32_t)(((s >> 18) ^ s) >> 27);
  uint32_t r = (uint32_t)(s >> 59);
  return (x >> r) | (x << ((32u - r) & 31u));
}
__device__ __forceinline__ float u01(uint64_t& s) { return (pcg(s) >> 8) * (1.0f / 16777216.0f); }

__global__ void k_sg_init(void* params, float* __restrict__ tab, uint64_t seed, float init_range) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= kTabEntries) return;
  uint64_t s = seed ^ (0x9e3779b97f4a7c15ull * (uint64_t)(e + 1));
  pcg(s);
  const float a = (2.0f * u01(s) - 1.0f) * init_range, b = (2.0f * u01(s) - 1.0f) * init_range;
  tab[2 * e] = a;
  tab[2 * e + 1] = b;
  ((__half2*)((char*)params + kOffTab16))[e] = __floats2half2_rn(a, b);
}

__global__ void k_sg_init_mlp(void* params, float* __restrict__ mlp, uint64_t seed) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNMlp) return;
  uint64_t s = seed ^ (0xd1b54a32d192ed03ull * (uint64_t)(p + 7));
  pcg(s);
  float r;
  if (p < kOffW2) r = sqrtf(6.0f / (kIn + kH));
  else if (p < kOffW3) r = sqrtf(6.0f / (kH + kH));
  else if (p < kOffB3) r = sqrtf(6.0f / (kH + 1));
  else r = 0.0f;
  const float w = (2.0f * u01(s) - 1.0f) * r;
  mlp[p] = w;
  char* pb = (char*)params;
  if (p < kOffW2) ((short*)(pb + kOffW1b))[p - kOffW1] = f2bf(w);
  else if (p < kOffW3) ((short*)(pb + kOffW2b))[p - kOffW2] = f2bf(w);
  else if (p < kOffB3) ((float*)(pb + kOffW3f))[p - kOffW3] = w;
  else ((SgMeta*)pb)->b3 = w;
}

// Training samples: a fraction near_frac lies within `radius` of the surface
// (a uniformly chosen face, a uniform point on it, a uniform offset in the
// ball), the rest uniform in the box [lo, hi].
__global__ void k_sg_sample(const float* __restrict__ V, const int32_t* __restrict__ F, int64_t nF, int64_t n,
                            float near_frac, float radius, float lo0, float lo1, float lo2, float hi0, float hi1,
                            float hi2, uint64_t seed, float* __restrict__ P) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t s = seed ^ (0x9e3779b97f4a7c15ull * (uint64_t)(i + 1));
  pcg(s);
  float p[3];
  if (u01(s) < near_frac && nF > 0) {
    const int64_t f = (int64_t)(((uint64_t)pcg(s) << 32 | pcg(s)) % (uint64_t)nF);
    float r1 = sqrtf(u01(s)), r2 = u01(s);
    const float a = 1.0f - r1, b = r1 * (1.0f - r2), c = r1 * r2;
    float dir[3], len2;
    do {
      for (int d = 0; d < 3; ++d) dir[d] = 2.0f * u01(s) - 1.0f;
      len2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
    } while (len2 > 1.0f || len2 < 1e-12f);
    const float rr = radius * cbrtf(u01(s)) * rsqrtf(len2);
    for (int d = 0; d < 3; ++d)
      p[d] = a * V[(int64_t)F[f * 3] * 3 + d] + b * V[(int64_t)F[f * 3 + 1] * 3 + d] +
             c * V[(int64_t)F[f * 3 + 2] * 3 + d] + dir[d] * rr;
  } else {
    const float lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
    for (int d = 0; d < 3; ++d) p[d] = lo[d] + (hi[d] - lo[d]) * u01(s);
  }
  for (int d = 0; d < 3; ++d) P[i * 3 + d] = p[d];
}

__global__ void k_iota_u32(uint32_t* p, int64_t n) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) p[t] = (uint32_t)t;
}

inline unsigned gsz(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace
}  // namespace a3d

using namespace a3d;

extern "C" int a3d_sg_sizes(int64_t* params_bytes, int64_t* state_bytes) {
  if (!params_bytes || !state_bytes) return fail(A3D_ERR_ARG, "a3d_sg_sizes: null argument");
  *params_bytes = kParamsBytes;
  *state_bytes = carve_s(nullptr).bytes;
  return A3D_OK;
}

extern "C" int a3d_sg_init(void* params, void* state, const float* lo_host, const float* hi_host, uint64_t seed,
                           a3d_stream_t stream) {
  if (!params || !state || !lo_host || !hi_host) return fail(A3D_ERR_ARG, "a3d_sg_init: null argument");
  SgMeta M;
  memset(&M, 0, sizeof(M));
  M.magic = (int32_t)kSgMagic;
  M.levels = kL;
  M.log2_t = kLog2T;
  M.features = 2;
  for (int d = 0; d < 3; ++d) {
    if (!(hi_host[d] > lo_host[d])) return fail(A3D_ERR_ARG, "a3d_sg_init: empty domain on axis %d", d);
    M.lo[d] = lo_host[d];
    M.inv_ext[d] = 1.0f / (hi_host[d] - lo_host[d]);
  }
  // N_l = floor(16 * b^l), b = (2048 / 16)^(1 / 15)
  const double b = pow(2048.0 / 16.0, 1.0 / (kL - 1));
  for (int l = 0; l < kL; ++l) {
    const int N = (int)floor(16.0 * pow(b, l) + 1e-9);
    M.scale[l] = (float)N;
    M.res1[l] = N + 1;
    M.dense[l] = (int64_t)(N + 1) * (N + 1) * (N + 1) <= kT ? 1 : 0;
  }
  hipStream_t s = (hipStream_t)stream;
  SgState st = carve_s(state);
  A3D_HIP_CHECK(hipMemcpyAsync(params, &M, sizeof(M), hipMemcpyHostToDevice, s));
  A3D_HIP_CHECK(hipMemsetAsync(state, 0, st.bytes, s));
  hipLaunchKernelGGL(k_sg_init, dim3(gsz(kTabEntries)), dim3(256), 0, s, params, st.tab, seed, 1e-4f);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sg_init_mlp, dim3(gsz(kNMlp)), dim3(256), 0, s, params, st.mlp, seed + 1);
  A3D_LAUNCH_CHECK();
  A3D_HIP_CHECK(hipStreamSynchronize(s));  // M lives on this stack frame
  return A3D_OK;
}

// Gradients of the mean squared error over one batch: hash-table gradient
// accumulated into the state, MLP gradient (and the loss at index 6209)
// reduced into the state's gradient vector.  grad_out (optional, device,
// kStride floats) receives a copy of the latter.
extern "C" int a3d_sg_grad(const void* params, void* state, const float* P, const float* target, int64_t B,
                           float* grad_out, a3d_stream_t stream) {
  if (!params || !state) return fail(A3D_ERR_ARG, "a3d_sg_grad: null argument");
  if (B <= 0 || B > (int64_t(1) << 28) || !P || !target) return fail(A3D_ERR_ARG, "a3d_sg_grad: bad batch");
  hipStream_t s = (hipStream_t)stream;
  SgState st = carve_s(state);
  const int64_t ntiles = (B + 31) / 32;
  int64_t blocks = (ntiles + kWavesS - 1) / kWavesS;
  if (blocks > kMaxTrainWaves / kWavesS) blocks = kMaxTrainWaves / kWavesS;
  hipLaunchKernelGGL(k_sg_train, dim3((unsigned)blocks), dim3(kBlkS), 0, s, params, st.tab, P, target, B,
                     1.0f / (float)B, st.tg, st.part);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sg_reduce, dim3(gsz(kOffLoss + 1, 64)), dim3(1024), 0, s, st.part, blocks * kWavesS, st.mg);
  A3D_LAUNCH_CHECK();
  if (grad_out) A3D_HIP_CHECK(hipMemcpyAsync(grad_out, st.mg, kStride * 4, hipMemcpyDeviceToDevice, s));
  return A3D_OK;
}

// Copy of the accumulated hash-table gradient (float [16 * 2^19 * 2]).
extern "C" int a3d_sg_table_grad(const void* state, float* out, a3d_stream_t stream) {
  if (!state || !out) return fail(A3D_ERR_ARG, "a3d_sg_table_grad: null argument");
  SgState st = carve_s((void*)state);
  A3D_HIP_CHECK(hipMemcpyAsync(out, st.tg, kTabEntries * 2 * 4, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return A3D_OK;
}

// One Adam step (bias-corrected, step t >= 1) on the accumulated gradients;
// clears the table gradient.
extern "C" int a3d_sg_apply(void* params, void* state, float lr, float beta1, float beta2, float eps, int64_t t,
                            a3d_stream_t stream) {
  if (!params || !state || t < 1) return fail(A3D_ERR_ARG, "a3d_sg_apply: bad argument");
  hipStream_t s = (hipStream_t)stream;
  SgState st = carve_s(state);
  AdamArgs a;
  a.b1 = beta1;
  a.b2 = beta2;
  a.eps = eps;
  a.lr_t = (float)(lr * sqrt(1.0 - pow((double)beta2, (double)t)) / (1.0 - pow((double)beta1, (double)t)));
  hipLaunchKernelGGL(k_sg_adam_table, dim3(gsz(kTabEntries)), dim3(256), 0, s, params, st.tab, st.tm, st.tv, st.tg,
                     a);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sg_adam_mlp, dim3(gsz(kNMlp)), dim3(256), 0, s, params, st.mlp, st.mm, st.mv, st.mg, a);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_sg_sample(const float* V, const int32_t* F, int64_t nF, int64_t n, float near_frac, float radius,
                             const float* lo_host, const float* hi_host, uint64_t seed, float* P,
                             a3d_stream_t stream) {
  if (n < 0 || !P || !lo_host || !hi_host || (nF > 0 && (!V || !F))) return fail(A3D_ERR_ARG, "a3d_sg_sample: bad argument");
  if (n == 0) return A3D_OK;
  hipLaunchKernelGGL(k_sg_sample, dim3(gsz(n)), dim3(256), 0, (hipStream_t)stream, V, F, nF, n, near_frac, radius,
                     lo_host[0], lo_host[1], lo_host[2], hi_host[0], hi_host[1], hi_host[2], seed, P);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_sg_infer(const void* params, const float* P, int64_t K, const uint32_t* perm, float* out,
                            a3d_stream_t stream) {
  if (!params || K < 0 || (K > 0 && (!P || !out))) return fail(A3D_ERR_ARG, "a3d_sg_infer: bad argument");
  if (perm && K > 0xffffffffll) return fail(A3D_ERR_ARG, "a3d_sg_infer: perm needs K < 2^32");
  if (K == 0) return A3D_OK;
  const int64_t ntiles = (K + 31) / 32;
  int64_t blocks = (ntiles + kWavesS - 1) / kWavesS;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_sg_infer, dim3((unsigned)blocks), dim3(kBlkS), 0, (hipStream_t)stream, params, P, perm, K, out,
                     0.0f, (uint8_t*)nullptr, (int32_t*)nullptr, (float*)nullptr, (float*)nullptr);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_sg_route_ws_bytes(int64_t K, int64_t* bytes) {
  if (K < 0 || !bytes) return fail(A3D_ERR_ARG, "a3d_sg_route_ws_bytes: bad argument");
  size_t t = 0;
  (void)rocprim::select(nullptr, t, (uint32_t*)nullptr, (uint8_t*)nullptr, (uint32_t*)nullptr,
                                      (int64_t*)nullptr, (int)(K > 1 ? K : 2));
  *bytes = al256(K > 0 ? K : 1) + al256(4 * (K > 0 ? K : 1)) + al256(4 * (K > 0 ? K : 1)) + 256 + al256((int64_t)t);
  return A3D_OK;
}

// Router: the network answers every query (dist[q]); queries predicted
// below `band` are re-answered exactly by the BVH wave walk (dist, face,
// closest, uvw); the others get face -1, closest = the query, uvw 0.
// perm (optional): a query order for coherence (a3d_query_order).  The count
// of exact queries is written to *n_near_host (synchronises the stream).
extern "C" int a3d_bvh_udf(const void* blob, const float* P, int64_t K, const uint32_t* perm, float* dist,
                           int32_t* face, float* closest, float* uvw, int flags, int64_t* counters,
                           a3d_stream_t stream);

extern "C" int a3d_sg_udf(const void* params, const void* blob, const float* P, int64_t K, const uint32_t* perm,
                          float band, float* dist, int32_t* face, float* closest, float* uvw, void* ws,
                          int64_t ws_bytes, int64_t* n_near_host, a3d_stream_t stream) {
  if (!params || !blob || K < 0 || !n_near_host) return fail(A3D_ERR_ARG, "a3d_sg_udf: bad argument");
  if (K > 0x7fffffffll) return fail(A3D_ERR_ARG, "a3d_sg_udf: K must be < 2^31");
  *n_near_host = 0;
  if (K == 0) return A3D_OK;
  if (!P || !dist || !face || !closest || !ws) return fail(A3D_ERR_ARG, "a3d_sg_udf: null input/output");
  int64_t need = 0;
  a3d_sg_route_ws_bytes(K, &need);
  if (ws_bytes < need) return fail(A3D_ERR_CAPACITY, "a3d_sg_udf: workspace %lld < %lld", (long long)ws_bytes, (long long)need);
  hipStream_t s = (hipStream_t)stream;
  char* w = (char*)ws;
  uint8_t* near = (uint8_t*)w;
  w += al256(K);
  uint32_t* ord = (uint32_t*)w;
  w += al256(4 * K);
  uint32_t* list = (uint32_t*)w;
  w += al256(4 * K);
  int64_t* cnt = (int64_t*)w;
  w += 256;
  size_t tb = (size_t)(need - (w - (char*)ws));
  const uint32_t* order = perm;
  if (!order) {
    hipLaunchKernelGGL(k_iota_u32, dim3(gsz(K)), dim3(256), 0, s, ord, K);
    A3D_LAUNCH_CHECK();
    order = ord;
  }
  const int64_t ntiles = (K + 31) / 32;
  int64_t blocks = (ntiles + kWavesS - 1) / kWavesS;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_sg_infer, dim3((unsigned)blocks), dim3(kBlkS), 0, s, params, P, order, K, dist, band, near,
                     face, closest, uvw);
  A3D_LAUNCH_CHECK();
  A3D_HIP_CHECK(rocprim::select(w, tb, order, near, list, cnt, (int)K, s));
  int64_t n_near = 0;
  A3D_HIP_CHECK(hipMemcpyAsync(&n_near, cnt, 8, hipMemcpyDeviceToHost, s));
  A3D_HIP_CHECK(hipStreamSynchronize(s));
  *n_near_host = n_near;
  if (n_near == 0) return A3D_OK;
  return a3d_bvh_udf(blob, P, n_near, list, dist, face, closest, uvw, 0, nullptr, stream);
}
