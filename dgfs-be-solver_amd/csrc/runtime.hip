// Stream capture into HIP graphs for the time steppers: a whole RK step
// (collision operator, advection, boundary updates and stage combinations,
// a few hundred launches) is recorded once on the solver's stream and replayed
// with one launch per step.  The captured kernels keep the device pointers and
// scalars of the recorded step, so a replay repeats exactly that step.
#include "common.hpp"

namespace dgfs {

struct GraphExec {
  int magic = 0x67726631;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static GraphExec* as_graph(void* p) {
  auto* g = static_cast<GraphExec*>(p);
  if (!g || g->magic != 0x67726631) return nullptr;
  return g;
}

}  // namespace dgfs

using namespace dgfs;

extern "C" int dgfs_graph_begin(void* stream) {
  if (!stream) return fail(DGFS_ERR_INVALID_VALUE, "capture needs a non-default stream");
  DGFS_HIP_TRY(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeRelaxed));
  return DGFS_OK;
}

extern "C" int dgfs_graph_end(void* stream, void** exec) {
  if (!stream || !exec) return fail(DGFS_ERR_INVALID_VALUE, "null stream or output");
  *exec = nullptr;
  hipGraph_t graph = nullptr;
  DGFS_HIP_TRY(hipStreamEndCapture(static_cast<hipStream_t>(stream), &graph));
  if (!graph) return fail(DGFS_ERR_HIP, "stream capture produced no graph");
  auto* g = new GraphExec();
  g->graph = graph;
  const hipError_t e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    delete g;
    return fail(DGFS_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  }
  *exec = g;
  return DGFS_OK;
}

extern "C" int dgfs_graph_launch(void* p, void* stream) {
  GraphExec* g = as_graph(p);
  if (!g) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid graph handle");
  DGFS_HIP_TRY(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
  return DGFS_OK;
}

extern "C" int dgfs_graph_nodes(void* p, long* count) {
  GraphExec* g = as_graph(p);
  if (!g) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid graph handle");
  if (!count) return fail(DGFS_ERR_INVALID_VALUE, "null output");
  size_t n = 0;
  DGFS_HIP_TRY(hipGraphGetNodes(g->graph, nullptr, &n));
  *count = static_cast<long>(n);
  return DGFS_OK;
}

extern "C" int dgfs_graph_destroy(void* p) {
  GraphExec* g = as_graph(p);
  if (!g) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid graph handle");
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  g->magic = 0;
  delete g;
  return DGFS_OK;
}
