// HIP graphs of plan executions: a fixed workload captured once and replayed
// with one launch call (no per-kernel host work).  No counterpart in the
// reference, which has no device code; the plans' execute calls allocate
// nothing, synchronise nothing and record no memset (the permanent queue
// cursor is reset by the finalize kernel, the SLOS vacuum coefficient is
// written by a kernel), which is what capture requires.
#include "common.h"
#include "../../include/pcvl_hip.h"

#include <new>
#include <string>

struct pcvl_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace pcvl {

// Captures enqueue(stream) on a private stream (default priority) into an
// instantiated graph.
template <class F>
static int capture(F enqueue, pcvl_graph *out)
{
    *out = nullptr;
    hipStream_t s = nullptr;
    PCVL_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(s);
        set_error(std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
        return ERR_HIP;
    }
    const int rc = enqueue(s);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    (void)hipStreamDestroy(s);
    if (rc != OK) {   // the enqueue's own error stays the reported one
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        return ERR_HIP;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        return ERR_HIP;
    }
    auto *G = new (std::nothrow) pcvl_graph_s();
    if (!G) {
        (void)hipGraphExecDestroy(x);
        (void)hipGraphDestroy(g);
        set_error("host allocation failed");
        return ERR_ALLOC;
    }
    G->graph = g;
    G->exec = x;
    *out = G;
    return OK;
}

} // namespace pcvl

using namespace pcvl;

extern "C" int pcvl_perm_graph_create(pcvl_perm_plan plan, const double *d_mats, double *d_out, int repeats,
                                      pcvl_graph *out)
{
    PCVL_ARG_CHECK(out != nullptr, "out graph pointer is null");
    PCVL_ARG_CHECK(plan != nullptr && d_mats && d_out, "null plan or device buffer");
    PCVL_ARG_CHECK(repeats >= 1 && repeats <= 1024, "repeats outside [1, 1024]");
    return capture([&](hipStream_t s) {
        for (int r = 0; r < repeats; r++) {
            const int rc = pcvl_perm_plan_execute(plan, d_mats, d_out, s);
            if (rc != OK) return rc;
        }
        return (int)OK;
    }, out);
}

extern "C" int pcvl_slos_graph_create(pcvl_slos_plan plan, const double *d_U, int mode, double *d_out, void *d_work,
                                      size_t work_bytes, int repeats, pcvl_graph *out)
{
    PCVL_ARG_CHECK(out != nullptr, "out graph pointer is null");
    PCVL_ARG_CHECK(plan != nullptr && d_U && d_out, "null plan or device buffer");
    PCVL_ARG_CHECK(repeats >= 1 && repeats <= 1024, "repeats outside [1, 1024]");
    return capture([&](hipStream_t s) {
        for (int r = 0; r < repeats; r++) {
            const int rc = pcvl_slos_plan_execute(plan, d_U, mode, d_out, d_work, work_bytes, s);
            if (rc != OK) return rc;
        }
        return (int)OK;
    }, out);
}

extern "C" int pcvl_graph_launch(pcvl_graph graph, hipStream_t stream)
{
    PCVL_ARG_CHECK(graph != nullptr && graph->exec != nullptr, "graph is null");
    PCVL_HIP_CHECK(hipGraphLaunch(graph->exec, stream));
    return OK;
}

extern "C" int pcvl_graph_destroy(pcvl_graph graph)
{
    if (!graph) return OK;
    hipError_t e1 = graph->exec ? hipGraphExecDestroy(graph->exec) : hipSuccess;
    hipError_t e2 = graph->graph ? hipGraphDestroy(graph->graph) : hipSuccess;
    delete graph;
    if (e1 != hipSuccess || e2 != hipSuccess) {
        set_error(std::string("destroying a graph: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
        return ERR_HIP;
    }
    return OK;
}
