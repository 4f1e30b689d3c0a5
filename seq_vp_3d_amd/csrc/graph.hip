// HIP graph capture of device-only segments of the solver, and the asynchronous copies those segments end with.
//
// A segment is a fixed sequence of library launches whose arguments are device addresses (a step length the host
// chooses travels through a page-locked host word that the segment's first node copies to the device). The host
// captures it once per set of buffer addresses and then launches the instantiated graph: one launch instead of a
// dozen, and no host work between the kernels of the segment. Capture is thread-local, so other host threads keep
// launching, allocating and synchronising on their own streams while one thread captures.
#include "abi_util.h"

extern "C" int seqvp_graph_begin(void* stream) {
  if (!stream) return SEQVP_EINVAL;   // the null stream cannot be captured
  return hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal) == hipSuccess ? SEQVP_OK
                                                                                                  : SEQVP_ELAUNCH;
}

extern "C" int seqvp_graph_end(void* stream, void** exec) {
  if (!stream || !exec) return SEQVP_EINVAL;
  *exec = nullptr;
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture((hipStream_t)stream, &graph);
  // a launch that failed while capturing invalidates the capture: report it after the capture is closed
  const hipError_t launch = hipGetLastError();
  if (e != hipSuccess || launch != hipSuccess || !graph) {
    if (graph) hipGraphDestroy(graph);
    return SEQVP_ELAUNCH;
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t ie = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ie != hipSuccess || !ex) return SEQVP_ELAUNCH;
  *exec = (void*)ex;
  return SEQVP_OK;
}

extern "C" int seqvp_graph_launch(void* exec, void* stream) {
  if (!exec) return SEQVP_EINVAL;
  return hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream) == hipSuccess ? SEQVP_OK : SEQVP_ELAUNCH;
}

extern "C" int seqvp_graph_destroy(void* exec) {
  if (!exec) return SEQVP_OK;
  return hipGraphExecDestroy((hipGraphExec_t)exec) == hipSuccess ? SEQVP_OK : SEQVP_ELAUNCH;
}

extern "C" int seqvp_copy_async(void* dst, const void* src, size_t bytes, void* stream) {
  if ((!dst || !src) && bytes) return SEQVP_EINVAL;
  if (!bytes) return SEQVP_OK;
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, (hipStream_t)stream) == hipSuccess ? SEQVP_OK
                                                                                             : SEQVP_ELAUNCH;
}
