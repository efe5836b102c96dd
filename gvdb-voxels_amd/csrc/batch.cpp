// batch.cpp -- see batch.h.
#include "batch.h"
#include "graph.h"
#include "../../include/gvdb_amd.h"

namespace gvdb {

Exec Batch::exec() {
    if (cap_) return Exec{stream, &cap_->arena, dsmall_, hsmall_, cap_};
    return Exec{stream, &arena_, dsmall_, hsmall_};
}

Batch::Batch(int dev, int mg) : device(dev), max_grids(mg) {}

Batch::~Batch() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    arena_.release();
    if (dsmall_) (void)hipFree(dsmall_);
    if (hsmall_) (void)hipHostFree(hsmall_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    shared.reset();   // the stream lives on while grids still run on it
}

int Batch::init() {
    if (max_grids < 1) { err = "batch: max_grids must be positive"; return GVDB_ERR_INVALID; }
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) {
        shared = SharedStream::make(device, e);
        stream = shared->s;
    }
    if (e == hipSuccess) e = arena_.init(size_t(4) << 20);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&dsmall_), size_t(max_grids) * SMALL_WORDS * 4);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&hsmall_), size_t(max_grids) * SMALL_WORDS * 4, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreate(&ev0_);
    if (e == hipSuccess) e = hipEventCreate(&ev1_);
    if (e != hipSuccess) {
        err = std::string("batch init: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? GVDB_ERR_OOM : GVDB_ERR_HIP;
    }
    return GVDB_OK;
}

int Batch::enter(Volume* const* v, int G, int* failed) {
    if (failed) *failed = -1;
    if (G < 1 || G > max_grids || !v) {
        err = "batch: grid count must be 1..max_grids";
        return GVDB_ERR_INVALID;
    }
    (void)hipSetDevice(device);
    for (int g = 0; g < G; g++) {
        if (!v[g] || v[g]->device != device) {
            if (failed) *failed = g;
            err = "batch: every volume must live on the batch's device";
            return GVDB_ERR_INVALID;
        }
        for (int k = 0; k < g; k++)
            if (v[k] == v[g]) {
                if (failed) *failed = g;
                err = "batch: a volume appears twice in one call";
                return GVDB_ERR_INVALID;
            }
        if (cap_) {
            // a recording cannot wait on another stream: the grid must already
            // run on the batch's stream (any earlier batched call moves it)
            if (v[g]->stream != stream || (v[g]->capturing() && v[g]->capture() != cap_)) {
                if (failed) *failed = g;
                err = "batch capture: every grid must take part in a batched call before the recording";
                if (!cap_->failed) {
                    cap_->failed = true;
                    cap_->err = err;
                }
                return GVDB_ERR_STATE;
            }
            v[g]->set_capture(cap_);
            continue;
        }
        if (v[g]->stream != stream) {
            int rc = v[g]->use_stream(shared);
            if (rc) {
                if (failed) *failed = g;
                err = v[g]->err;
                return rc;
            }
        }
    }
    if (!started_ && !cap_) {
        (void)hipEventRecord(ev0_, stream);
        started_ = true;
    }
    return GVDB_OK;
}

int Batch::leave(int rc, Volume* const* v, int G, int* failed) {
    if (!cap_) (void)hipEventRecord(ev1_, stream);
    else if (rc && !cap_->failed) {
        cap_->failed = true;
        cap_->err = (failed && *failed >= 0 && *failed < G) ? v[*failed]->err : err;
    }
    if (rc && failed && *failed >= 0 && *failed < G) err = v[*failed]->err;
    return rc;
}

#define BATCH_CALL(expr)                               \
    int f_ = -1;                                       \
    int* fp_ = failed ? failed : &f_;                  \
    int rc_ = enter(v, G, fp_);                        \
    if (rc_) return rc_;                               \
    rc_ = (expr);                                      \
    return leave(rc_, v, G, fp_)

int Batch::rebuild_topology(Volume* const* v, int G, const int* n, const float* radius, const float* orig3, int* failed) {
    BATCH_CALL(Volume::multi_rebuild_topology(v, G, n, radius, orig3, exec(), fp_));
}
int Batch::finish_topology(Volume* const* v, int G, bool commit, bool bounds, int* failed) {
    BATCH_CALL(Volume::multi_finish_topology(v, G, commit, bounds, exec(), fp_));
}
int Batch::update_atlas(Volume* const* v, int G, int* failed) {
    BATCH_CALL(Volume::multi_update_atlas(v, G, exec(), fp_));
}
int Batch::clear_channel(Volume* const* v, int G, int chan, int* failed) {
    BATCH_CALL(Volume::multi_clear_channel(v, G, chan, exec(), fp_));
}
int Batch::insert_points_subcell(Volume* const* v, int G, int scs, const int* n, const float* radius,
                                 const float* trans3, int* len, int* failed) {
    BATCH_CALL(Volume::multi_insert(v, G, scs, n, radius, trans3, len, exec(), fp_));
}
int Batch::gather_density(Volume* const* v, int G, int scs, const int* len, const float* radius, int chan, bool accum,
                          int* failed) {
    BATCH_CALL(Volume::multi_gather(v, G, scs, len, radius, chan, accum, false, exec(), fp_));
}
int Batch::gather_level_set(Volume* const* v, int G, int scs, const int* len, const float* radius, int chan, bool accum,
                            int* failed) {
    BATCH_CALL(Volume::multi_gather(v, G, scs, len, radius, chan, accum, true, exec(), fp_));
}
int Batch::update_apron(Volume* const* v, int G, int chan, float boundval, int* failed) {
    BATCH_CALL(Volume::multi_apron(v, G, chan, boundval, exec(), fp_));
}
int Batch::render(Volume* const* v, int G, const int* shading, const int* chan, const int* buf, int* failed) {
    BATCH_CALL(Volume::multi_render(v, G, shading, chan, buf, exec(), fp_));
}

int Batch::synchronize() {
    if (cap_) {
        err = "batch: synchronize while recording a graph";
        return GVDB_ERR_STATE;
    }
    hipError_t e = hipStreamSynchronize(stream);
    arena_.reset();
    if (e != hipSuccess) {
        err = std::string("batch synchronize: ") + hipGetErrorString(e);
        return GVDB_ERR_HIP;
    }
    return GVDB_OK;
}

int Batch::elapsed_ms(float* ms) {
    *ms = 0.f;
    if (!started_ || cap_) return GVDB_OK;
    hipError_t e = hipEventSynchronize(ev1_);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, ev0_, ev1_);
    started_ = false;
    arena_.reset();
    if (e != hipSuccess) {
        err = std::string("batch elapsed: ") + hipGetErrorString(e);
        return GVDB_ERR_HIP;
    }
    return GVDB_OK;
}

}  // namespace gvdb
