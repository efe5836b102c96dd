// graph.cpp -- see graph.h.
#include "graph.h"
#include "batch.h"
#include "../../include/gvdb_amd.h"
#include <algorithm>
#include <atomic>

namespace gvdb {

static std::atomic<unsigned long long> g_alloc_gen{0};
unsigned long long alloc_generation() { return g_alloc_gen.load(); }
void bump_alloc_generation() { g_alloc_gen.fetch_add(1); }

// Argument tables of the multi-grid launches recorded in one graph.
static const size_t GRAPH_ARENA_BYTES = size_t(4) << 20;

Graph::~Graph() {
    (void)hipSetDevice(device);
    if (stream && pending) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    arena.release();
    for (Volume* v : vols) v->forget_graph(this);
}

void Graph::note(Volume* v) {
    if (std::find(vols.begin(), vols.end(), v) != vols.end()) return;
    vols.push_back(v);
    topo.push_back(v->topo_generation());
    v->remember_graph(this);
}

static int start(Graph* g, int device, hipStream_t s, std::string& err) {
    g->device = device;
    g->stream = s;
    hipError_t e = g->arena.init(GRAPH_ARENA_BYTES);
    g->arena.fixed = true;
    g->alloc_gen = alloc_generation();
    if (e == hipSuccess) {
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        e = hipStreamIsCapturing(s, &st);
        if (e == hipSuccess && st != hipStreamCaptureStatusNone) {
            err = "capture: the stream is already capturing";
            return GVDB_ERR_STATE;
        }
    }
    // relaxed: the host-side work of the recorded calls (argument staging,
    // pinned readback slots) is plain host code, not stream work
    if (e == hipSuccess) e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
        err = std::string("capture begin: ") + hipGetErrorString(e);
        return GVDB_ERR_HIP;
    }
    return GVDB_OK;
}

int capture_finish(Graph* g, hipStream_t s, std::string& err) {
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s, &graph);
    g->graph = graph;
    if (e != hipSuccess) {
        err = std::string("capture end: ") + hipGetErrorString(e);
        return GVDB_ERR_HIP;
    }
    if (g->failed) {
        err = "capture: a recorded call failed: " + g->err;
        return GVDB_ERR_STATE;
    }
    if (g->alloc_gen != alloc_generation()) {
        err = "capture: device buffers were allocated while recording; run the chain once without capture first";
        return GVDB_ERR_STATE;
    }
    if (!graph) {
        err = "capture: nothing was recorded";
        return GVDB_ERR_STATE;
    }
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        err = std::string("graph instantiate: ") + hipGetErrorString(e);
        return GVDB_ERR_HIP;
    }
    return GVDB_OK;
}

int capture_begin(Volume& v, Graph*& out, std::string& err) {
    out = nullptr;
    if (v.capturing()) {
        err = "capture: this grid is already recording";
        return GVDB_ERR_STATE;
    }
    (void)hipSetDevice(v.device);
    std::unique_ptr<Graph> g(new Graph());
    g->keep = v.stream_owner();
    int rc = start(g.get(), v.device, v.stream, err);
    if (rc) return rc;
    g->note(&v);
    v.set_capture(g.get());
    out = g.release();
    return GVDB_OK;
}

int capture_end(Volume& v, Graph* g, std::string& err) {
    if (!v.capturing() || v.capture() != g) {
        err = "capture: this grid is not recording";
        return GVDB_ERR_STATE;
    }
    v.set_capture(nullptr);
    return capture_finish(g, v.stream, err);
}

int graph_launch(Graph* g, std::string& err) {
    if (!g->exec) {
        err = "graph: not instantiated";
        return GVDB_ERR_STATE;
    }
    if (g->failed) {
        err = "graph: " + g->err;
        return GVDB_ERR_STATE;
    }
    if (g->alloc_gen != alloc_generation()) {
        err = "graph: device buffers were reallocated since the capture (capture again)";
        return GVDB_ERR_STATE;
    }
    for (size_t k = 0; k < g->vols.size(); k++)
        if (g->vols[k]->topo_generation() != g->topo[k] || g->vols[k]->capturing()) {
            err = "graph: a grid's topology changed since the capture, or it is recording (capture again)";
            return GVDB_ERR_STATE;
        }
    (void)hipSetDevice(g->device);
    hipError_t e = hipGraphLaunch(g->exec, g->stream);
    if (e != hipSuccess) {
        err = std::string("graph launch: ") + hipGetErrorString(e);
        return GVDB_ERR_HIP;
    }
    g->pending = true;
    return GVDB_OK;
}

int graph_synchronize(Graph* g, std::string& err) {
    (void)hipSetDevice(g->device);
    hipError_t e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) {
        err = std::string("graph synchronize: ") + hipGetErrorString(e);
        return GVDB_ERR_HIP;
    }
    if (!g->pending) return GVDB_OK;
    g->pending = false;
    // the recorded inserts' readbacks: entry counts, and points the
    // bucketed insert could not place
    int rc = GVDB_OK;
    for (const Graph::Insert& in : g->inserts) {
        int r = in.v->finish_recorded_insert(in.slot);
        if (r != GVDB_OK && rc == GVDB_OK) {
            rc = r;
            err = in.v->err;
        }
    }
    return rc;
}

// ------------------------------------------------------------ batches ----
int Batch::capture_begin(Graph*& out) {
    out = nullptr;
    if (cap_) {
        err = "capture: this batch is already recording";
        return GVDB_ERR_STATE;
    }
    (void)hipSetDevice(device);
    // the recorded calls start from a drained stream and a fresh timing window
    int rc = synchronize();
    if (rc) return rc;
    std::unique_ptr<Graph> g(new Graph());
    g->keep = shared;
    rc = start(g.get(), device, stream, err);
    if (rc) return rc;
    cap_ = g.release();
    out = cap_;
    return GVDB_OK;
}

int Batch::capture_end(Graph* g) {
    if (!cap_ || cap_ != g) {
        err = "capture: this batch is not recording";
        return GVDB_ERR_STATE;
    }
    for (Volume* v : g->vols) v->set_capture(nullptr);
    cap_ = nullptr;
    return capture_finish(g, stream, err);
}

}  // namespace gvdb
