// graph.h -- HIP graphs of the per-frame chain ClearChannel ->
// InsertPointsSubcell -> GatherDensity -> UpdateApron -> Render.
//
// Between capture_begin and capture_end, the calls on one grid (or one
// batch) are recorded into a HIP graph instead of run; a launch replays them
// with one hipGraphLaunch.  The chain needs no host round trip once the
// topology is built: the sizes the stages read (bricks, subcells, buckets)
// are fixed by the topology, and the one value the insert hands back (its
// subcell entry count) stays on the device and reaches pinned host memory
// through a copy recorded in the graph, read by graph_synchronize.
//
// A replay runs the recorded launches with their recorded arguments (point
// count, radius, camera) over the current contents of the device buffers
// (points, atlas, render buffers).  It is refused when the topology of a
// grid changed or any library buffer was reallocated since the capture; the
// insert's points must still fall inside the topology (a replay whose insert
// met a point outside every brick reports GVDB_ERR_STATE at
// graph_synchronize, since the bucketed insert leaves such points out where
// the eager call falls back to the reference's per-entry binning).
#pragma once
#include <memory>
#include <string>
#include <vector>
#include "volume.h"

namespace gvdb {

// Bumped by every device allocation or free the library makes for a grid
// (pools, atlas, scratch); a graph records it at capture.
unsigned long long alloc_generation();
void bump_alloc_generation();

struct Graph {
    int device = 0;
    hipStream_t stream = nullptr;
    std::shared_ptr<SharedStream> keep;   // the stream it replays on stays alive with it
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    Arena arena;                          // argument tables of captured multi-grid launches
    std::vector<Volume*> vols;            // grids the capture recorded calls for
    std::vector<unsigned> topo;           // their topology generations at capture
    struct Insert {
        Volume* v;
        const uint32_t* slot;             // pinned words the recorded readback fills
    };
    std::vector<Insert> inserts;
    unsigned long long alloc_gen = 0;
    bool failed = false;                  // a recorded call failed: capture_end refuses the graph
    std::string err;
    bool pending = false;                 // launched, not yet synchronised
    ~Graph();
    void note(Volume* v);                 // v takes part in the capture
};

// Capture on a grid's stream, or on a batch's stream (batch.cpp).
int capture_begin(Volume& v, Graph*& out, std::string& err);
int capture_end(Volume& v, Graph* g, std::string& err);
int capture_finish(Graph* g, hipStream_t s, std::string& err);   // ends capture on s, instantiates
int graph_launch(Graph* g, std::string& err);
int graph_synchronize(Graph* g, std::string& err);

}  // namespace gvdb
