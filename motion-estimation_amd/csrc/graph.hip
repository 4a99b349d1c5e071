// Stream capture into HIP graphs: the batched reconstruction records one
// outer iteration of every volume (its CG iterations and LM update) once and
// replays it, so the host issues one launch per iteration instead of one per
// kernel.  Capture is thread-local: an allocation on this thread while
// capturing is an error rather than a silent hazard.
#include "me_internal.h"
#include "../../include/motion_estimation_amd.h"

using namespace me;

extern "C" {

int me_graph_begin(hipStream_t stream) {
  ME_TRY_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  return 0;
}

int me_graph_end(hipStream_t stream, void** exec) {
  ME_REQUIRE(exec, "graph_end: null exec");
  *exec = nullptr;
  hipGraph_t graph = nullptr;
  ME_TRY_HIP(hipStreamEndCapture(stream, &graph));
  hipGraphExec_t e = nullptr;
  const hipError_t rc = hipGraphInstantiate(&e, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  ME_TRY_HIP(rc);
  *exec = (void*)e;
  return 0;
}

int me_graph_abort(hipStream_t stream) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  ME_TRY_HIP(hipStreamIsCapturing(stream, &st));
  if (st != hipStreamCaptureStatusNone) {
    hipGraph_t graph = nullptr;
    (void)hipStreamEndCapture(stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
  }
  (void)hipGetLastError();
  return 0;
}

int me_graph_launch(void* exec, hipStream_t stream) {
  ME_REQUIRE(exec, "graph_launch: null exec");
  ME_TRY_HIP(hipGraphLaunch((hipGraphExec_t)exec, stream));
  return 0;
}

int me_graph_destroy(void* exec) {
  if (exec) ME_TRY_HIP(hipGraphExecDestroy((hipGraphExec_t)exec));
  return 0;
}

}  // extern "C"
