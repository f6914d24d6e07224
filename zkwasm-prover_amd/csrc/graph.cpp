// Capture and replay of per-round launch sequences as HIP graphs.
//
// A prover round issues the same sequence of launches (batched NTTs, batched
// MSMs with the device tail, element-wise kernels) with the same buffers for
// every proof of a shape.  zk_graph_begin / zk_graph_end record what the
// library enqueues on a stream into an executable graph; zk_graph_launch
// replays it with one call.  Only asynchronous entry points may run while
// capturing (ntt, zk_ntt_batch, zk_msm_batch_async, msm, field_op and the
// other kernels-only entries); the library clears memory with kernels, never
// with memset nodes.  The capturing stream must not be the legacy default
// stream.
#include <hip/hip_runtime.h>

#include "zkwasm_prover_kernel.h"

extern "C" {

hipError_t zk_graph_begin(hipStream_t stream) {
    if (!stream) return hipErrorInvalidValue;
    return hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
}

hipError_t zk_graph_end(hipStream_t stream, void **exec) {
    if (!stream || !exec) return hipErrorInvalidValue;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (e != hipSuccess) return e;
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return e;
    *exec = (void *)x;
    return hipSuccess;
}

hipError_t zk_graph_launch(void *exec, hipStream_t stream) {
    if (!exec) return hipErrorInvalidValue;
    return hipGraphLaunch((hipGraphExec_t)exec, stream);
}

hipError_t zk_graph_destroy(void *exec) {
    if (!exec) return hipSuccess;
    return hipGraphExecDestroy((hipGraphExec_t)exec);
}

}  // extern "C"
