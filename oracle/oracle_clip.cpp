// Oracle — CPU restatement of the reference's triangle/box polygon clip.
//
// TEST INFRASTRUCTURE ONLY (see oracle_bvh.cpp for the ground rules).
//
// Restated from Atom3D atom3d/kernels/cumtv_kernels.cu:
//   clip_with_plane<8>           :206-261  (Sutherland-Hodgman, keep dist <= eps)
//   sat_clip_polygon_kernel      :273-460  (+x,-x,+y,-y,+z,-z planes; centroid =
//                                 box-clamped vertex mean projected onto the
//                                 triangle with clamped barycentrics; fan area
//                                 around the clamped mean; area < eps -> 0)
// Pinning (tests/test_oracle_clip.py): against exact FP64 clipping (vertex
// sets, polygon area by the shoelace formula, centroid inside box and on the
// triangle) on random cases, and closed forms (triangle inside the box: the
// triangle itself; box inside the triangle's plane region: a rectangle).
#include <cmath>
#include <cstdint>

namespace {

constexpr int MAXV = 8;

int clip_plane(const float in[MAXV][3], int n_in, const float nrm[3], float d, float out[MAXV][3], float eps) {
  if (n_in <= 0) return 0;
  float dist[MAXV];
  bool inside[MAXV];
  for (int i = 0; i < n_in; ++i) {
    dist[i] = in[i][0] * nrm[0] + in[i][1] * nrm[1] + in[i][2] * nrm[2] - d;
    inside[i] = dist[i] <= eps;
  }
  int n = 0;
  for (int i = 0; i < n_in; ++i) {
    int j = (i + 1 == n_in) ? 0 : i + 1;
    const float* P = in[i];
    const float* Q = in[j];
    float dP = dist[i], dQ = dist[j];
    bool iP = inside[i], iQ = inside[j];
    if (iP && iQ) {
      out[n][0] = Q[0]; out[n][1] = Q[1]; out[n][2] = Q[2]; ++n;
    } else if (iP != iQ) {
      float den = dP - dQ;
      if (fabsf(den) > eps) {
        float t = dP / den;
        out[n][0] = P[0] + t * (Q[0] - P[0]);
        out[n][1] = P[1] + t * (Q[1] - P[1]);
        out[n][2] = P[2] + t * (Q[2] - P[2]);
        ++n;
      }
      if (!iP) {
        if (n >= MAXV) return MAXV;
        out[n][0] = Q[0]; out[n][1] = Q[1]; out[n][2] = Q[2]; ++n;
      }
    }
    if (n >= MAXV) return MAXV;
  }
  return n;
}

}  // namespace

extern "C" {

// One candidate per k: box a = cand_a[k] of (bmin, bmax), triangle t =
// cand_t[k] of tris9.  mode 0: hit only; 1: + centroid and area; 2: + polygon.
// Every output is written for every k (zeros where the reference leaves its
// zero-initialised buffers untouched).
void orc_sat_clip(const float* bmin, const float* bmax, const float* tris9, const int64_t* cand_a,
                  const int64_t* cand_t, int64_t N, int mode, float eps, uint8_t* hit, int32_t* counts,
                  float* poly, float* centroids, float* areas) {
  static const float nxp[3] = {1, 0, 0}, nyp[3] = {0, 1, 0}, nzp[3] = {0, 0, 1};
  static const float nxn[3] = {-1, 0, 0}, nyn[3] = {0, -1, 0}, nzn[3] = {0, 0, -1};
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < N; ++k) {
    const float* lo = bmin + cand_a[k] * 3;
    const float* hi = bmax + cand_a[k] * 3;
    const float* tv = tris9 + cand_t[k] * 9;
    float A[MAXV][3], B[MAXV][3];
    for (int v = 0; v < 3; ++v) for (int c = 0; c < 3; ++c) A[v][c] = tv[v * 3 + c];
    hit[k] = 0; counts[k] = 0; areas[k] = 0.0f;
    centroids[k * 3] = centroids[k * 3 + 1] = centroids[k * 3 + 2] = 0.0f;
    if (mode == 2 && poly) for (int i = 0; i < MAXV * 3; ++i) poly[k * MAXV * 3 + i] = 0.0f;
    int n = 3;
    n = clip_plane(A, n, nxp, hi[0], B, eps); if (!n) continue;
    n = clip_plane(B, n, nxn, -lo[0], A, eps); if (!n) continue;
    n = clip_plane(A, n, nyp, hi[1], B, eps); if (!n) continue;
    n = clip_plane(B, n, nyn, -lo[1], A, eps); if (!n) continue;
    n = clip_plane(A, n, nzp, hi[2], B, eps); if (!n) continue;
    n = clip_plane(B, n, nzn, -lo[2], A, eps); if (!n) continue;
    hit[k] = 1;
    int cnt = n > MAXV ? MAXV : n;
    counts[k] = cnt;
    if (mode == 2 && poly)
      for (int i = 0; i < cnt; ++i) for (int c = 0; c < 3; ++c) poly[(k * MAXV + i) * 3 + c] = A[i][c];
    float cx = 0, cy = 0, cz = 0;
    for (int i = 0; i < cnt; ++i) { cx += A[i][0]; cy += A[i][1]; cz += A[i][2]; }
    cx /= (float)cnt; cy /= (float)cnt; cz /= (float)cnt;
    cx = fminf(fmaxf(cx, lo[0]), hi[0]);
    cy = fminf(fmaxf(cy, lo[1]), hi[1]);
    cz = fminf(fmaxf(cz, lo[2]), hi[2]);
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
      centroids[k * 3] = ax; centroids[k * 3 + 1] = ay; centroids[k * 3 + 2] = az;
      continue;
    }
    float v = (d11 * d20 - d01 * d21) / den;
    float w = (d00 * d21 - d01 * d20) / den;
    float u = 1.0f - v - w;
    if (u < 0) u = 0;
    if (v < 0) v = 0;
    if (w < 0) w = 0;
    float nm = u + v + w;
    if (nm <= eps) { u = 1; v = 0; w = 0; nm = 1; }
    u /= nm; v /= nm; w /= nm;
    centroids[k * 3] = u * ax + v * bx + w * qx;
    centroids[k * 3 + 1] = u * ay + v * by + w * qy;
    centroids[k * 3 + 2] = u * az + v * bz + w * qz;
    float area = 0;
    for (int i = 0; i < cnt; ++i) {
      int j = (i + 1) % cnt;
      float ax_ = A[i][0] - cx, ay_ = A[i][1] - cy, az_ = A[i][2] - cz;
      float bx_ = A[j][0] - cx, by_ = A[j][1] - cy, bz_ = A[j][2] - cz;
      float crx = ay_ * bz_ - az_ * by_, cry = az_ * bx_ - ax_ * bz_, crz = ax_ * by_ - ay_ * bx_;
      area += 0.5f * sqrtf(crx * crx + cry * cry + crz * crz);
    }
    areas[k] = area < eps ? 0.0f : area;
  }
}

}  // extern "C"
