// Triangle/box polygon clip for gfx950 (intersect_aabb modes 2 and 3).
//
// Replaces sat_clip_polygon_kernel / sat_clip_polygon_cuda (Atom3D
// atom3d/kernels/cumtv_kernels.cu:193-460, 1035-1100): Sutherland-Hodgman
// clip of each candidate triangle against its box's six planes in the order
// +x, -x, +y, -y, +z, -z (a vertex is kept when its signed plane distance is
// <= eps; at most 8 vertices), then the box-clamped vertex mean projected on
// the triangle with clamped barycentrics, and the fan area around the clamped
// mean (areas below eps reported as 0).  Same FP32 operation order as the
// reference, so the outputs match the CPU oracle bit for bit.  One difference:
// the reference can write a 9th vertex past its 8-entry array before
// truncating to 8; here the 9th vertex is never written (the first 8 and the
// count are the same).
#include "a3d_common.hpp"

namespace a3d {
namespace {

constexpr int kBlk = 256;
constexpr int kMaxV = 8;

__device__ __forceinline__ int clip_plane(const float (&in)[kMaxV][3], int n_in, float n0, float n1, float n2, float d,
                                          float (&out)[kMaxV][3], float eps) {
  if (n_in <= 0) return 0;
  float dist[kMaxV];
  bool inside[kMaxV];
#pragma unroll
  for (int i = 0; i < kMaxV; ++i) {
    if (i < n_in) {
      dist[i] = in[i][0] * n0 + in[i][1] * n1 + in[i][2] * n2 - d;
      inside[i] = dist[i] <= eps;
    }
  }
  int n = 0;
#pragma unroll
  for (int i = 0; i < kMaxV; ++i) {
    if (i >= n_in) break;
    int j = (i + 1 == n_in) ? 0 : i + 1;
    float Px = in[i][0], Py = in[i][1], Pz = in[i][2];
    float Qx = in[j][0], Qy = in[j][1], Qz = in[j][2];
    float dP = dist[i], dQ = dist[j];
    bool iP = inside[i], iQ = inside[j];
    if (iP && iQ) {
      out[n][0] = Qx; out[n][1] = Qy; out[n][2] = Qz; ++n;
    } else if (iP != iQ) {
      float den = dP - dQ;
      if (fabsf(den) > eps) {
        float t = dP / den;
        out[n][0] = Px + t * (Qx - Px);
        out[n][1] = Py + t * (Qy - Py);
        out[n][2] = Pz + t * (Qz - Pz);
        ++n;
      }
      if (!iP) {
        if (n >= kMaxV) return kMaxV;
        out[n][0] = Qx; out[n][1] = Qy; out[n][2] = Qz; ++n;
      }
    }
    if (n >= kMaxV) return kMaxV;
  }
  return n;
}

__global__ __launch_bounds__(kBlk) void k_sat_clip(const float* __restrict__ bmin, const float* __restrict__ bmax,
                                                    const float* __restrict__ tris9,
                                                    const int64_t* __restrict__ cand_a,
                                                    const int64_t* __restrict__ cand_t, int64_t N, int mode, float eps,
                                                    uint8_t* __restrict__ hit, int32_t* __restrict__ counts,
                                                    float* __restrict__ poly, float* __restrict__ centroids,
                                                    float* __restrict__ areas) {
  int64_t k = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (k >= N) return;
  const float* lo = bmin + cand_a[k] * 3;
  const float* hi = bmax + cand_a[k] * 3;
  const float* tv = tris9 + cand_t[k] * 9;
  float A[kMaxV][3], B[kMaxV][3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    A[v][0] = tv[v * 3];
    A[v][1] = tv[v * 3 + 1];
    A[v][2] = tv[v * 3 + 2];
  }
  float lx = lo[0], ly = lo[1], lz = lo[2], hx = hi[0], hy = hi[1], hz = hi[2];
  int n = 3;
  n = clip_plane(A, n, 1.0f, 0.0f, 0.0f, hx, B, eps);
  if (n) n = clip_plane(B, n, -1.0f, 0.0f, 0.0f, -lx, A, eps);
  if (n) n = clip_plane(A, n, 0.0f, 1.0f, 0.0f, hy, B, eps);
  if (n) n = clip_plane(B, n, 0.0f, -1.0f, 0.0f, -ly, A, eps);
  if (n) n = clip_plane(A, n, 0.0f, 0.0f, 1.0f, hz, B, eps);
  if (n) n = clip_plane(B, n, 0.0f, 0.0f, -1.0f, -lz, A, eps);
  const int cnt = n > kMaxV ? kMaxV : n;
  hit[k] = cnt > 0 ? 1 : 0;
  counts[k] = cnt;
  if (mode == 2 && poly) {
#pragma unroll
    for (int i = 0; i < kMaxV; ++i) {
      bool on = i < cnt;
      poly[(k * kMaxV + i) * 3] = on ? A[i][0] : 0.0f;
      poly[(k * kMaxV + i) * 3 + 1] = on ? A[i][1] : 0.0f;
      poly[(k * kMaxV + i) * 3 + 2] = on ? A[i][2] : 0.0f;
    }
  }
  float ox = 0.0f, oy = 0.0f, oz = 0.0f, area_out = 0.0f;
  if (cnt > 0) {
    float cx = 0, cy = 0, cz = 0;
    for (int i = 0; i < cnt; ++i) { cx += A[i][0]; cy += A[i][1]; cz += A[i][2]; }
    cx /= (float)cnt; cy /= (float)cnt; cz /= (float)cnt;
    cx = fminf(fmaxf(cx, lx), hx);
    cy = fminf(fmaxf(cy, ly), hy);
    cz = fminf(fmaxf(cz, lz), hz);
    float ax = tv[0], ay = tv[1], az = tv[2], bx = tv[3], by = tv[4], bz = tv[5], qx = tv[6], qy = tv[7], qz = tv[8];
    float e1x = bx - ax, e1y = by - ay, e1z = bz - az;
    float e2x = qx - ax, e2y = qy - ay, e2z = qz - az;
    float vx = cx - ax, vy = cy - ay, vz = cz - az;
    float d00 = e1x * e1x + e1y * e1y + e1z * e1z;
    float d01 = e1x * e2x + e1y * e2y + e1z * e2z;
    float d11 = e2x * e2x + e2y * e2y + e2z * e2z;
    float d20 = vx * e1x + vy * e1y + vz * e1z;
    float d21 = vx * e2x + vy * e2y + vz * e2z;
    float den = d00 * d11 - d01 * d01;
    if (fabsf(den) < eps * eps) {
      ox = ax; oy = ay; oz = az;  // degenerate triangle: first vertex, area left 0
    } else {
      float v = (d11 * d20 - d01 * d21) / den;
      float w = (d00 * d21 - d01 * d20) / den;
      float u = 1.0f - v - w;
      if (u < 0) u = 0;
      if (v < 0) v = 0;
      if (w < 0) w = 0;
      float nm = u + v + w;
      if (nm <= eps) { u = 1; v = 0; w = 0; nm = 1; }
      u /= nm; v /= nm; w /= nm;
      ox = u * ax + v * bx + w * qx;
      oy = u * ay + v * by + w * qy;
      oz = u * az + v * bz + w * qz;
      float area = 0;
      for (int i = 0; i < cnt; ++i) {
        int j = (i + 1) % cnt;
        float ax_ = A[i][0] - cx, ay_ = A[i][1] - cy, az_ = A[i][2] - cz;
        float bx_ = A[j][0] - cx, by_ = A[j][1] - cy, bz_ = A[j][2] - cz;
        float crx = ay_ * bz_ - az_ * by_, cry = az_ * bx_ - ax_ * bz_, crz = ax_ * by_ - ay_ * bx_;
        area += 0.5f * sqrtf(crx * crx + cry * cry + crz * crz);
      }
      area_out = area < eps ? 0.0f : area;
    }
  }
  centroids[k * 3] = ox;
  centroids[k * 3 + 1] = oy;
  centroids[k * 3 + 2] = oz;
  areas[k] = area_out;
}

}  // namespace
}  // namespace a3d

using namespace a3d;

extern "C" int a3d_sat_clip(const float* bmin, const float* bmax, const float* tris9, const int64_t* cand_a,
                            const int64_t* cand_t, int64_t N, int mode, float eps, uint8_t* hit, int32_t* counts,
                            float* poly, float* centroids, float* areas, a3d_stream_t stream) {
  if (N < 0) return fail(A3D_ERR_ARG, "a3d_sat_clip: N < 0");
  if (mode < 0 || mode > 2) return fail(A3D_ERR_ARG, "a3d_sat_clip: mode=%d not in {0,1,2}", mode);
  if (N == 0) return A3D_OK;
  if (!bmin || !bmax || !tris9 || !cand_a || !cand_t || !hit || !counts || !centroids || !areas)
    return fail(A3D_ERR_ARG, "a3d_sat_clip: null input/output");
  if (mode == 2 && !poly) return fail(A3D_ERR_ARG, "a3d_sat_clip: mode 2 needs the polygon output");
  hipLaunchKernelGGL(k_sat_clip, dim3((unsigned)div_up(N, kBlk)), dim3(kBlk), 0, (hipStream_t)stream, bmin, bmax,
                     tris9, cand_a, cand_t, N, mode, eps, hit, counts, poly, centroids, areas);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}
