// batch.h -- many grids per launch: the batched form of the point path.
//
// A Batch owns one stream, argument staging and per-grid readback slots.
// Each stage call takes a list of volumes (grids of any size, each with its
// own pools and atlas) and runs that stage for all of them with one launch
// per kernel (Volume::multi_*).  Volumes taking part run on the batch's
// stream from their first batched call on.
#pragma once
#include <string>
#include <vector>
#include "volume.h"

namespace gvdb {

class Batch {
public:
    Batch(int device, int max_grids);
    ~Batch();
    int init();
    std::string err;
    int device;
    int max_grids;
    hipStream_t stream = nullptr;
    std::shared_ptr<SharedStream> shared;

    int rebuild_topology(Volume* const* v, int G, const int* n, const float* radius, const float* orig3, int* failed);
    int finish_topology(Volume* const* v, int G, bool commit, bool bounds, int* failed);
    int update_atlas(Volume* const* v, int G, int* failed);
    int clear_channel(Volume* const* v, int G, int chan, int* failed);
    int insert_points_subcell(Volume* const* v, int G, int scs, const int* n, const float* radius, const float* trans3,
                              int* len, int* failed);
    int gather_level_set(Volume* const* v, int G, int scs, const int* len, const float* radius, int chan, bool accum,
                         int* failed);
    int gather_density(Volume* const* v, int G, int scs, const int* len, const float* radius, int chan, bool accum,
                       int* failed);
    int update_apron(Volume* const* v, int G, int chan, float boundval, int* failed);
    int render(Volume* const* v, int G, const int* shading, const int* chan, const int* buf, int* failed);
    // HIP graph capture of the batched chain (graph.cpp): between these the
    // stage calls are recorded, not run
    int capture_begin(Graph*& out);
    int capture_end(Graph* g);
    int synchronize();
    // milliseconds between the start of the first and the end of the last
    // stage call since the previous synchronize (HIP events on the stream)
    int elapsed_ms(float* ms);

private:
    Arena arena_;
    uint32_t* dsmall_ = nullptr;
    uint32_t* hsmall_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    bool started_ = false;
    Graph* cap_ = nullptr;   // the graph being recorded, if any
    int enter(Volume* const* v, int G, int* failed);
    int leave(int rc, Volume* const* v, int G, int* failed);
    Exec exec();
};

}  // namespace gvdb
