/* gvdb_amd.hpp -- header-only C++ facade over the C ABI of gvdb_amd.h.
 *
 * Mirrors the point-path methods of nvdb::VolumeGVDB
 * (source/gvdb_library/src/gvdb_volume_gvdb.h:319-901 of NVIDIA GVDB Voxels
 * 1.1.1) with the same names, argument order and meaning, so a caller of the
 * reference can switch by changing the include and the namespace.  Like the
 * reference, the methods return void; a failed call records its status and
 * message (status(), error()), and with exceptions enabled (the default) also
 * throws gvdb_amd::Error.  Vector3DF/Vector3DI are replaced by the small
 * structs below with the same x, y, z members.
 */
#ifndef GVDB_AMD_HPP
#define GVDB_AMD_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "gvdb_amd.h"

namespace gvdb_amd {

struct Vector3DF {
    float x, y, z;
    Vector3DF(float a = 0, float b = 0, float c = 0) : x(a), y(b), z(c) {}
};
struct Vector3DI {
    int x, y, z;
    Vector3DI(int a = 0, int b = 0, int c = 0) : x(a), y(b), z(c) {}
};

/* DataPtr (gvdb_allocator.h:35-56) reduced to what the point path reads:
 * a device pointer, element count, byte stride and byte offset. */
struct DataPtr {
    const void* gpu = nullptr;
    int num = 0;
    int stride = 12;
    int subdim_offset = 0;
};

enum { T_UCHAR4 = GVDB_T_UCHAR4, T_FLOAT = GVDB_T_FLOAT, T_HALF = GVDB_T_HALF };
enum { SHADE_TRILINEAR = GVDB_SHADE_TRILINEAR, SHADE_VOLUME = GVDB_SHADE_VOLUME };

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

class VolumeGVDB {
public:
    explicit VolumeGVDB(int device = 0, bool throw_on_error = true) : throw_(throw_on_error) {
        check(gvdb_create(&h_, device), "gvdb_create");
    }
    ~VolumeGVDB() {
        if (h_) gvdb_destroy(h_);
    }
    VolumeGVDB(const VolumeGVDB&) = delete;
    VolumeGVDB& operator=(const VolumeGVDB&) = delete;

    gvdb_handle handle() const { return h_; }
    int status() const { return status_; }
    std::string error() const { return h_ ? gvdb_last_error(h_) : "no volume"; }

    /* Configure(q4, q3, q2, q1, q0): brick last, leaf-first log2 dims {q0..q4} */
    void Configure(int q4, int q3, int q2, int q1, int q0) {
        int r[5] = {q0, q1, q2, q3, q4};
        check(gvdb_configure(h_, 5, r, nullptr), "Configure");
    }
    void Configure(int levs, const int* r, const int* numcnt) { check(gvdb_configure(h_, levs, r, numcnt), "Configure"); }
    void SetChannelDefault(int cx, int cy, int cz) { check(gvdb_set_channel_default(h_, cx, cy, cz), "SetChannelDefault"); }
    void AddChannel(unsigned char chan, int dt, int apron, int filter = 1, int border = 1,
                    Vector3DI axiscnt = Vector3DI(0, 0, 0)) {
        check(gvdb_add_channel(h_, chan, dt, apron, filter, border, axiscnt.x, axiscnt.y, axiscnt.z), "AddChannel");
    }
    void SetEpsilon(float eps, int maxiter) { check(gvdb_set_epsilon(h_, eps, maxiter), "SetEpsilon"); }

    /* SetPoints(pos, vel, clr): positions, and the colours (uint32 per point) when clr.gpu is set; the velocities
     * are not read on this path */
    void SetPoints(const DataPtr& pos, const DataPtr& = DataPtr(), const DataPtr& clr = DataPtr()) {
        check(gvdb_set_points_gpu(h_, pos.gpu, pos.num, pos.subdim_offset, pos.stride), "SetPoints");
        check(gvdb_set_colors_gpu(h_, clr.gpu, clr.gpu ? clr.num : 0, clr.subdim_offset, clr.stride ? clr.stride : 4),
              "SetPoints");
    }
    void SetPointsHost(const float* xyz, int n) { check(gvdb_upload_points(h_, xyz, n), "SetPoints"); }
    void SetColorChannel(unsigned char chan) { check(gvdb_set_color_channel(h_, chan), "SetColorChannel"); }
    void SetTransform(Vector3DF pretrans, Vector3DF scal, Vector3DF angs, Vector3DF trans) {
        const float p[3] = {pretrans.x, pretrans.y, pretrans.z}, s[3] = {scal.x, scal.y, scal.z},
                    a[3] = {angs.x, angs.y, angs.z}, t[3] = {trans.x, trans.y, trans.z};
        check(gvdb_set_transform(h_, p, s, a, t), "SetTransform");
    }

    void RequestFullRebuild(bool on) { check(gvdb_request_full_rebuild(h_, on ? 1 : 0), "RequestFullRebuild"); }
    void RebuildTopology(int num_pnts, float radius, Vector3DF orig) {
        const float o[3] = {orig.x, orig.y, orig.z};
        check(gvdb_rebuild_topology(h_, num_pnts, radius, o), "RebuildTopology");
    }
    void AccumulateTopology(int num_pnts, float radius, Vector3DF orig, int iDepth = 1) {
        const float o[3] = {orig.x, orig.y, orig.z};
        check(gvdb_accumulate_topology(h_, num_pnts, radius, o, iDepth), "AccumulateTopology");
    }
    void RebuildTopologyCPU(int num_pnts, Vector3DF orig, const float* host_xyz) {
        const float o[3] = {orig.x, orig.y, orig.z};
        check(gvdb_rebuild_topology_cpu(h_, num_pnts, o, host_xyz), "RebuildTopologyCPU");
    }
    void ActivateSpace(Vector3DF pos) {
        const float p[3] = {pos.x, pos.y, pos.z};
        check(gvdb_activate_space(h_, p), "ActivateSpace");
    }
    void Clear() { check(gvdb_clear(h_), "Clear"); }
    void FinishTopology(bool pCommitPool = true, bool pComputeBounds = true) {
        check(gvdb_finish_topology(h_, pCommitPool, pComputeBounds), "FinishTopology");
    }
    void UpdateAtlas() { check(gvdb_update_atlas(h_), "UpdateAtlas"); }
    void ClearChannel(unsigned char chan) { check(gvdb_clear_channel(h_, chan), "ClearChannel"); }

    void InsertPointsSubcell(int subcell_size, int num_pnts, float pRadius, Vector3DF trans, int& pSCPntsLength) {
        const float t[3] = {trans.x, trans.y, trans.z};
        check(gvdb_insert_points_subcell(h_, subcell_size, num_pnts, pRadius, t, &pSCPntsLength), "InsertPointsSubcell");
    }
    void GatherDensity(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength,
                       int chanDensity, int chanClr, bool bAccum = false) {
        const float t[3] = {trans.x, trans.y, trans.z};
        check(gvdb_gather_density(h_, subcell_size, num_pnts, radius, t, pSCPntsLength, chanDensity, chanClr,
                                  bAccum ? 1 : 0), "GatherDensity");
    }
    void GatherLevelSet(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength,
                        int chanLevelset, int chanClr, bool bAccum = false) {
        const float t[3] = {trans.x, trans.y, trans.z};
        check(gvdb_gather_level_set(h_, subcell_size, num_pnts, radius, t, pSCPntsLength, chanLevelset, chanClr,
                                    bAccum ? 1 : 0), "GatherLevelSet");
    }
    void InsertPointsSubcell_FP16(int subcell_size, int num_pnts, float pRadius, Vector3DF trans, int& pSCPntsLength) {
        const float t[3] = {trans.x, trans.y, trans.z};
        check(gvdb_insert_points_subcell_fp16(h_, subcell_size, num_pnts, pRadius, t, &pSCPntsLength),
              "InsertPointsSubcell_FP16");
    }
    void GatherLevelSet_FP16(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength,
                             int chanLevelset, int chanClr) {
        const float t[3] = {trans.x, trans.y, trans.z};
        check(gvdb_gather_level_set_fp16(h_, subcell_size, num_pnts, radius, t, pSCPntsLength, chanLevelset, chanClr),
              "GatherLevelSet_FP16");
    }
    void UpdateApron(unsigned char chan, float boundval = 0.0f) { check(gvdb_update_apron(h_, chan, boundval), "UpdateApron"); }
    void Compute(int effect, unsigned char channel, int num_iterations, Vector3DF parameters, bool bUpdateApron,
                 bool skipOverAprons, float boundval = 0.0f) {
        const float p[3] = {parameters.x, parameters.y, parameters.z};
        check(gvdb_compute(h_, effect, channel, num_iterations, p, bUpdateApron ? 1 : 0, skipOverAprons ? 1 : 0,
                           boundval), "Compute");
    }
    void ComputeKernel(void* user_module, void* user_kernel, unsigned char channel, bool bUpdateApron,
                       bool skipOverAprons) {
        (void)user_module;   // the function handle carries its module
        check(gvdb_compute_kernel(h_, user_kernel, channel, bUpdateApron ? 1 : 0, skipOverAprons ? 1 : 0),
              "ComputeKernel");
    }
    void RenderKernel(void* user_kernel, unsigned char in_channel = 0, unsigned char outbuf = 0) {
        check(gvdb_render_kernel(h_, user_kernel, in_channel, outbuf), "RenderKernel");
    }
    void SaveVBX(const std::string& fname) { check(gvdb_save_vbx(h_, fname.c_str()), "SaveVBX"); }
    bool LoadVBX(const std::string& fname, int force_maj = 0, int force_min = 0) {
        const int rc = gvdb_load_vbx(h_, fname.c_str(), force_maj, force_min);
        check(rc, "LoadVBX");
        return rc == GVDB_OK;
    }

    /* Scene setters of Scene / Camera3D used by the point samples */
    void SetCameraOrbit(Vector3DF angs, Vector3DF to, float dist, float dolly, float fov = 40.f, float znear = 0.1f,
                        float zfar = 5000.f) {
        const float a[3] = {angs.x, angs.y, angs.z}, t[3] = {to.x, to.y, to.z};
        check(gvdb_set_camera_orbit(h_, a, t, dist, dolly, fov, znear, zfar), "SetCamera");
    }
    void SetLight(Vector3DF pos) {
        const float p[3] = {pos.x, pos.y, pos.z};
        check(gvdb_set_light(h_, p), "SetLight");
    }
    void SetBackgroundClr(float r, float g, float b, float a) {
        const float c[4] = {r, g, b, a};
        check(gvdb_set_background(h_, c), "SetBackgroundClr");
    }
    void SetExtinct(float a, float b, float c) { check(gvdb_set_extinct(h_, a, b, c), "SetExtinct"); }
    void SetSteps(float directstep, float shadowstep, float fineStep) { check(gvdb_set_steps(h_, directstep, shadowstep, fineStep), "SetSteps"); }
    void SetCutoff(float a, float b, float c) { check(gvdb_set_cutoff(h_, a, b, c), "SetCutoff"); }
    void SetVolumeRange(float thresh, float vmin, float vmax) { check(gvdb_set_volume_range(h_, thresh, vmin, vmax), "SetVolumeRange"); }
    void SetShadowParams(float amount, float bias, float c) { check(gvdb_set_shadow_params(h_, amount, bias, c), "SetShadowParams"); }
    void LinearTransferFunc(float t0, float t1, const float a[4], const float b[4]) {
        check(gvdb_linear_transfer_func(h_, t0, t1, a, b), "LinearTransferFunc");
    }
    void CommitTransferFunc() { check(gvdb_commit_transfer_func(h_), "CommitTransferFunc"); }
    void AddRenderBuf(int chan, int width, int height, int byteperpix) {
        check(gvdb_add_render_buf(h_, chan, width, height, byteperpix), "AddRenderBuf");
    }
    void Render(char shading, unsigned char chan, unsigned char rbuf) { check(gvdb_render(h_, shading, chan, rbuf), "Render"); }
    void ReadRenderBuf(int chan, unsigned char* outptr) { check(gvdb_read_render_buf(h_, chan, outptr), "ReadRenderBuf"); }

    /* Node access: getNumTotalNodes / getNodeAtLevel over a fetched pool */
    uint64_t getNumTotalNodes(int lev) {
        uint64_t n = 0, s = 0;
        check(gvdb_get_pool(h_, 0, lev, nullptr, &n, &s), "getNumTotalNodes");
        return n;
    }
    std::vector<gvdb_node> getNodes(int lev) {
        uint64_t n = 0, s = 0;
        check(gvdb_get_pool(h_, 0, lev, nullptr, &n, &s), "getNodes");
        std::vector<gvdb_node> out(n);
        if (n) check(gvdb_get_pool(h_, 0, lev, out.data(), &n, &s), "getNodes");
        return out;
    }
    void* getStream() {
        void* s = nullptr;
        check(gvdb_get_stream(h_, &s), "getStream");
        return s;
    }
    void Synchronize() { check(gvdb_synchronize(h_), "Synchronize"); }

    /* Render autotuner and setter (no counterpart in the reference) */
    int TuneRender(int shading, int chan = 0, int buf = 0, float* ms3 = nullptr) {
        int best = 0;
        check(gvdb_tune_render(h_, shading, chan, buf, &best, ms3), "TuneRender");
        return best;
    }
    static int SetRenderUnroll(int shading, int unroll) { return gvdb_set_render_unroll(shading, unroll); }

private:
    void check(int rc, const char* what) {
        status_ = rc;
        if (rc != GVDB_OK && throw_) throw Error(rc, std::string(what) + ": " + error());
    }
    gvdb_handle h_ = nullptr;
    int status_ = GVDB_OK;
    bool throw_ = true;
};

/* Many grids per launch (gvdb_batch_*; no counterpart in the reference):
 * each method runs one stage of the point path over `vols` with one kernel
 * launch per kernel of the stage.  Per-grid arguments are vectors with one
 * entry per volume.  A failure throws (or records) with the failing grid's
 * index in failed(). */
class GridBatch {
public:
    explicit GridBatch(int device = 0, int max_grids = 256, bool throw_on_error = true) : throw_(throw_on_error) {
        check(gvdb_batch_create(&b_, device, max_grids), "gvdb_batch_create");
    }
    ~GridBatch() {
        if (b_) gvdb_batch_destroy(b_);
    }
    GridBatch(const GridBatch&) = delete;
    GridBatch& operator=(const GridBatch&) = delete;
    int status() const { return status_; }
    int failed() const { return failed_; }

    void RebuildTopology(const std::vector<VolumeGVDB*>& vols, const std::vector<int>& n,
                         const std::vector<float>& radius) {
        auto hs = handles(vols);
        check(gvdb_batch_rebuild_topology(b_, int(hs.size()), hs.data(), n.data(), radius.data(), nullptr, &failed_),
              "RebuildTopology");
    }
    void FinishTopology(const std::vector<VolumeGVDB*>& vols, bool commit = false, bool bounds = true) {
        auto hs = handles(vols);
        check(gvdb_batch_finish_topology(b_, int(hs.size()), hs.data(), commit, bounds, &failed_), "FinishTopology");
    }
    void UpdateAtlas(const std::vector<VolumeGVDB*>& vols) {
        auto hs = handles(vols);
        check(gvdb_batch_update_atlas(b_, int(hs.size()), hs.data(), &failed_), "UpdateAtlas");
    }
    void ClearChannel(const std::vector<VolumeGVDB*>& vols, int chan) {
        auto hs = handles(vols);
        check(gvdb_batch_clear_channel(b_, int(hs.size()), hs.data(), chan, &failed_), "ClearChannel");
    }
    std::vector<int> InsertPointsSubcell(const std::vector<VolumeGVDB*>& vols, int subcell_size,
                                         const std::vector<int>& n, const std::vector<float>& radius) {
        auto hs = handles(vols);
        std::vector<int> len(hs.size(), 0);
        check(gvdb_batch_insert_points_subcell(b_, int(hs.size()), hs.data(), subcell_size, n.data(), radius.data(),
                                               nullptr, len.data(), &failed_),
              "InsertPointsSubcell");
        return len;
    }
    void GatherDensity(const std::vector<VolumeGVDB*>& vols, int subcell_size, const std::vector<int>& len,
                       const std::vector<float>& radius, int chan_density, bool accumulate = false) {
        auto hs = handles(vols);
        check(gvdb_batch_gather_density(b_, int(hs.size()), hs.data(), subcell_size, len.data(), radius.data(),
                                        chan_density, accumulate, &failed_),
              "GatherDensity");
    }
    void GatherLevelSet(const std::vector<VolumeGVDB*>& vols, int subcell_size, const std::vector<int>& len,
                        const std::vector<float>& radius, int chan_levelset, bool accumulate = false) {
        auto hs = handles(vols);
        check(gvdb_batch_gather_level_set(b_, int(hs.size()), hs.data(), subcell_size, len.data(), radius.data(),
                                          chan_levelset, accumulate, &failed_),
              "GatherLevelSet");
    }
    void UpdateApron(const std::vector<VolumeGVDB*>& vols, int chan, float boundval) {
        auto hs = handles(vols);
        check(gvdb_batch_update_apron(b_, int(hs.size()), hs.data(), chan, boundval, &failed_), "UpdateApron");
    }
    void Render(const std::vector<VolumeGVDB*>& vols, int shading, int chan, int buf) {
        auto hs = handles(vols);
        std::vector<int> sh(hs.size(), shading), ch(hs.size(), chan), bf(hs.size(), buf);
        check(gvdb_batch_render(b_, int(hs.size()), hs.data(), sh.data(), ch.data(), bf.data(), &failed_), "Render");
    }
    void Synchronize() { check(gvdb_batch_synchronize(b_), "Synchronize"); }

private:
    static std::vector<gvdb_handle> handles(const std::vector<VolumeGVDB*>& vols) {
        std::vector<gvdb_handle> hs;
        for (auto* v : vols) hs.push_back(v->handle());
        return hs;
    }
    void check(int rc, const char* what) {
        status_ = rc;
        if (rc != GVDB_OK && throw_)
            throw Error(rc, std::string(what) + ": " + (b_ ? gvdb_batch_last_error(b_) : "no batch"));
    }
    gvdb_batch b_ = nullptr;
    int status_ = GVDB_OK;
    int failed_ = -1;
    bool throw_ = true;
};

/* Serving loop over resident grids (gvdb_server_*; no counterpart in the
 * reference); see gvdb_amd.h for the request semantics. */
class Server {
public:
    enum { PLAIN = 0, CONTINUOUS = 1 };
    Server(int device, unsigned long long hbm_budget, int max_batch, bool throw_on_error = true)
        : throw_(throw_on_error) {
        check(gvdb_server_create(&s_, device, hbm_budget, max_batch), "gvdb_server_create");
    }
    ~Server() {
        if (s_) gvdb_server_destroy(s_);
    }
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;
    int status() const { return status_; }

    int AddGrid(VolumeGVDB& v, const float* host_xyz, int n, float radius) {
        return check(gvdb_server_add_grid(s_, v.handle(), host_xyz, n, radius), "AddGrid");
    }
    void Reset() { check(gvdb_server_reset(s_), "Reset"); }
    int SubmitRender(int grid, double arrival, int w, int h, int shading, const float angs[3], float dist,
                     unsigned char* out_rgba) {
        return check(gvdb_server_submit_render(s_, grid, arrival, w, h, shading, angs, dist, out_rgba),
                     "SubmitRender");
    }
    int SubmitInsert(int grid, double arrival, const float* host_xyz, int n) {
        return check(gvdb_server_submit_insert(s_, grid, arrival, host_xyz, n), "SubmitInsert");
    }
    void ClearRequests() { check(gvdb_server_clear_requests(s_), "ClearRequests"); }
    /* latency/status: one entry per submitted request; returns the wall time (s) */
    double Run(int mode, std::vector<double>& latency, std::vector<int>& status, size_t requests) {
        latency.assign(requests, 0.0);
        status.assign(requests, 0);
        double wall = 0.0;
        check(gvdb_server_run(s_, mode, latency.data(), status.data(), &wall), "Run");
        return wall;
    }

private:
    int check(int rc, const char* what) {
        status_ = rc < 0 ? rc : GVDB_OK;
        if (rc < 0 && throw_)
            throw Error(rc, std::string(what) + ": " + (s_ ? gvdb_server_last_error(s_) : "no server"));
        return rc;
    }
    gvdb_server s_ = nullptr;
    int status_ = GVDB_OK;
    bool throw_ = true;
};

}  // namespace gvdb_amd

#endif
