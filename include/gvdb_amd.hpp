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

enum { T_FLOAT = GVDB_T_FLOAT, T_HALF = GVDB_T_HALF };
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

    /* SetPoints(pos, vel, clr): only the positions are read on this path */
    void SetPoints(const DataPtr& pos, const DataPtr& = DataPtr(), const DataPtr& = DataPtr()) {
        check(gvdb_set_points_gpu(h_, pos.gpu, pos.num, pos.subdim_offset, pos.stride), "SetPoints");
    }
    void SetPointsHost(const float* xyz, int n) { check(gvdb_upload_points(h_, xyz, n), "SetPoints"); }

    void RequestFullRebuild(bool on) { check(gvdb_request_full_rebuild(h_, on ? 1 : 0), "RequestFullRebuild"); }
    void RebuildTopology(int num_pnts, float radius, Vector3DF orig) {
        const float o[3] = {orig.x, orig.y, orig.z};
        check(gvdb_rebuild_topology(h_, num_pnts, radius, o), "RebuildTopology");
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
    void UpdateApron(unsigned char chan, float boundval = 0.0f) { check(gvdb_update_apron(h_, chan, boundval), "UpdateApron"); }

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

private:
    void check(int rc, const char* what) {
        status_ = rc;
        if (rc != GVDB_OK && throw_) throw Error(rc, std::string(what) + ": " + error());
    }
    gvdb_handle h_ = nullptr;
    int status_ = GVDB_OK;
    bool throw_ = true;
};

}  // namespace gvdb_amd

#endif
