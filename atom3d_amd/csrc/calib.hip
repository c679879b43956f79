// Counter calibration for the VALU roofline: a kernel whose only work is
// dependency-free FP32 fused multiply-adds, so that SQ_INSTS_VALU over its
// cycles (GRBM_GUI_ACTIVE) measures the rate at which the chip issues plain
// 32-bit VALU instructions.  The walks' own issue rates are read against it.
// No counterpart in the reference.
#include "a3d_common.hpp"

using namespace a3d;

namespace {

constexpr int kCalibBlk = 256;

__global__ __launch_bounds__(kCalibBlk) void k_calib_valu(float* __restrict__ out, int iters, float s) {
  float a[8];
  for (int j = 0; j < 8; ++j) a[j] = (float)threadIdx.x * 1e-3f + (float)j;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = fmaf(a[j], s, 0.5f);
  }
  float r = 0.0f;
  for (int j = 0; j < 8; ++j) r += a[j];
  out[(int64_t)blockIdx.x * kCalibBlk + threadIdx.x] = r;
}

}  // namespace

extern "C" int a3d_calib_valu(float* out, int64_t n, int iters, a3d_stream_t stream) {
  if (!out || n <= 0 || n % kCalibBlk || n / kCalibBlk > 0x7fffffff || iters < 0)
    return fail(A3D_ERR_ARG, "a3d_calib_valu: n=%lld must be a positive multiple of %d, iters=%d >= 0", (long long)n,
                kCalibBlk, iters);
  k_calib_valu<<<(unsigned)(n / kCalibBlk), kCalibBlk, 0, (hipStream_t)stream>>>(out, iters, 0.999f);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}
