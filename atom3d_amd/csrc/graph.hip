// HIP-graph capture helpers: record the work a caller queues on a stream
// between a3d_graph_begin and a3d_graph_end into an instantiated graph and
// replay it with a3d_graph_launch.  No counterpart in the reference.
#include "a3d_common.hpp"

using namespace a3d;

extern "C" int a3d_graph_begin(a3d_stream_t stream) {
  A3D_HIP_CHECK(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed));
  return A3D_OK;
}

extern "C" int a3d_graph_end(a3d_stream_t stream, int64_t* handle) {
  if (!handle) return fail(A3D_ERR_ARG, "a3d_graph_end: null handle");
  hipGraph_t g = nullptr;
  A3D_HIP_CHECK(hipStreamEndCapture((hipStream_t)stream, &g));
  hipGraphExec_t e = nullptr;
  hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (err != hipSuccess) return fail(A3D_ERR_HIP, "a3d_graph_end: hipGraphInstantiate: %s", hipGetErrorString(err));
  *handle = (int64_t)(intptr_t)e;
  return A3D_OK;
}

extern "C" int a3d_graph_launch(int64_t handle, a3d_stream_t stream) {
  if (!handle) return fail(A3D_ERR_ARG, "a3d_graph_launch: null handle");
  A3D_HIP_CHECK(hipGraphLaunch((hipGraphExec_t)(intptr_t)handle, (hipStream_t)stream));
  return A3D_OK;
}

extern "C" int a3d_graph_destroy(int64_t handle) {
  if (handle) A3D_HIP_CHECK(hipGraphExecDestroy((hipGraphExec_t)(intptr_t)handle));
  return A3D_OK;
}
