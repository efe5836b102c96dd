// serve.h -- a serving loop over many resident grids.
//
// Grids stay resident in HBM.  Requests (renders of any size with their own
// camera; point inserts that append points to a grid and rebuild it:
// topology, atlas, subcell insert, gather, apron) arrive over time.  Two
// schedules run the same requests:
//   plain       one request at a time in arrival order through the one-grid
//               calls (what a caller of the reference would do);
//   continuous  every pass admits all requests that have arrived, as long as
//               their grid has nothing in flight (per-grid order is kept),
//               batch slots are free and the HBM budget holds; admitted
//               renders go out as one batched launch on the render stream
//               while admitted inserts run as one batched pipeline on the
//               insert stream, so the two overlap on the device, and
//               requests arriving meanwhile join the next pass.
// Latency = completion - arrival, per request.
#pragma once
#include <memory>
#include <string>
#include <vector>
#include "batch.h"
#include "volume.h"

namespace gvdb {

class Server {
public:
    Server(int device, unsigned long long budget_bytes, int max_batch);
    ~Server();
    int init();
    std::string err;

    // A resident grid: configured volume (channel 0 present), its initial
    // points (copied into a server-owned growing buffer) and the splat radius.
    int add_grid(Volume* v, const float* host_xyz, int n, float radius);
    // Builds every grid from its initial points (batched) and clears requests'
    // results; grids return to their initial points.
    int reset();

    int submit_render(int grid, double arrival, int w, int h, int shading, const float angs[3], float dist,
                      unsigned char* out_rgba);
    int submit_insert(int grid, double arrival, const float* host_xyz, int n);
    void clear_requests();

    // mode 0 plain, 1 continuous.  latency/status: one entry per submitted
    // request, in submission order.  wall: first arrival to last completion.
    int run(int mode, double* latency, int* status, double* wall);

    unsigned long long device_bytes() const;

private:
    struct Grid {
        Volume* v = nullptr;
        DevBuf pts;            // packed float3, grows
        int n0 = 0, n = 0;
        std::vector<float> init_host;
        float radius = 2.f;
        float centre[3] = {0, 0, 0};
        float extent = 1.f;
    };
    struct Req {
        int kind = 0;          // 0 render, 1 insert
        int grid = 0;
        double arrival = 0;
        int w = 0, h = 0, shading = SHADE_VOLUME;
        float angs[3] = {0, 0, 0};
        float dist = 0;
        unsigned char* out = nullptr;
        const float* pts = nullptr;
        int n = 0;
        double done = -1;
        int status = 0;
        int slot = -1;
    };
    int device_;
    unsigned long long budget_;
    int max_batch_;
    std::vector<Grid> grids_;
    std::vector<Req> reqs_;
    std::unique_ptr<Batch> ib_, rb_;      // insert and render batches (own streams)
    std::vector<unsigned char*> stage_;   // pinned readback slots, 1024^2 RGBA each
    hipEvent_t rev_ = nullptr;

    int append(Grid& g, const float* host, int n, hipStream_t s);
    int build(const std::vector<int>& which, Batch& b);
    int render_setup(Req& r);
    int run_plain(double t0);
    int run_continuous(double t0);
};

}  // namespace gvdb
