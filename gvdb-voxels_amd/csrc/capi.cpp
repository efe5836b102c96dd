// capi.cpp -- extern "C" entry points of include/gvdb_amd.h over gvdb::Volume.
#include "volume.h"
#include "kernels.h"
#include "batch.h"
#include "serve.h"
#include "graph.h"
#include "../../include/gvdb_amd.h"
#include <new>
#include <cstring>
#include <vector>

struct gvdb_volume {
    gvdb::Volume v;
    explicit gvdb_volume(int d) : v(d) {}
};

struct gvdb_batch_s {
    gvdb::Batch b;
    gvdb_batch_s(int d, int m) : b(d, m) {}
};

struct gvdb_server_s {
    gvdb::Server s;
    gvdb_server_s(int d, unsigned long long b, int m) : s(d, b, m) {}
};

struct gvdb_graph_s {
    std::unique_ptr<gvdb::Graph> g;
    std::string err;
};

static_assert(sizeof(gvdb_node) == sizeof(gvdb::Node), "gvdb_node mirrors the 64-byte node");
static_assert(GVDB_KNOB_COUNT == gvdb::KNOB_COUNT, "the header names every launch-shape knob");

#define H_OR_FAIL(h)                     \
    if (!(h)) return GVDB_ERR_INVALID;   \
    (void)hipSetDevice((h)->v.device)

// calls outside the per-frame chain fail (and fail the recording) while the
// grid records a graph
#define NOT_RECORDING(h, what) \
    if ((h)->v.capturing()) return (h)->v.not_recordable(what)

extern "C" {

int gvdb_create(gvdb_handle* out, int device) {
    if (!out) return GVDB_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return GVDB_ERR_HIP;
    gvdb_volume* h = new (std::nothrow) gvdb_volume(device);
    if (!h) return GVDB_ERR_OOM;
    int rc = h->v.init();
    if (rc) { delete h; return rc; }
    *out = h;
    return GVDB_OK;
}

int gvdb_destroy(gvdb_handle h) {
    if (!h) return GVDB_ERR_INVALID;
    delete h;
    return GVDB_OK;
}

const char* gvdb_last_error(gvdb_handle h) { return h ? h->v.err.c_str() : "null handle"; }

int gvdb_configure(gvdb_handle h, int levels, const int* log2dims, const int* initcnt) {
    H_OR_FAIL(h); NOT_RECORDING(h, "Configure");
    if (!log2dims) return GVDB_ERR_INVALID;
    return h->v.configure(levels, log2dims, initcnt);
}
int gvdb_set_channel_default(gvdb_handle h, int cx, int cy, int cz) { H_OR_FAIL(h); NOT_RECORDING(h, "SetChannelDefault"); return h->v.set_channel_default(cx, cy, cz); }
int gvdb_add_channel(gvdb_handle h, int chan, int dtype, int apron, int filter, int border, int ax, int ay, int az) {
    H_OR_FAIL(h); NOT_RECORDING(h, "AddChannel");
    return h->v.add_channel(chan, dtype, apron, filter, border, ax, ay, az);
}
int gvdb_set_epsilon(gvdb_handle h, float eps, int maxiter) { H_OR_FAIL(h); return h->v.set_epsilon(eps, maxiter); }
int gvdb_set_points_gpu(gvdb_handle h, const void* dptr, int n, int offset, int stride) {
    H_OR_FAIL(h); NOT_RECORDING(h, "SetPoints");
    return h->v.set_points_gpu(dptr, n, offset, stride);
}
int gvdb_upload_points(gvdb_handle h, const float* host_xyz, int n) { H_OR_FAIL(h); NOT_RECORDING(h, "SetPoints"); return h->v.upload_points(host_xyz, n); }
int gvdb_set_colors_gpu(gvdb_handle h, const void* dptr, int n, int offset, int stride) {
    H_OR_FAIL(h); NOT_RECORDING(h, "SetPoints");
    return h->v.set_colors_gpu(dptr, n, offset, stride);
}
int gvdb_upload_colors(gvdb_handle h, const uint32_t* host_rgba, int n) {
    H_OR_FAIL(h); NOT_RECORDING(h, "SetPoints");
    return h->v.upload_colors(host_rgba, n);
}
int gvdb_set_color_channel(gvdb_handle h, int chan) {
    H_OR_FAIL(h);
    return h->v.set_color_channel(chan);
}
int gvdb_request_full_rebuild(gvdb_handle h, int on) { H_OR_FAIL(h); return h->v.request_full_rebuild(on != 0); }
int gvdb_rebuild_topology(gvdb_handle h, int n, float radius, const float orig[3]) {
    H_OR_FAIL(h); NOT_RECORDING(h, "RebuildTopology");
    const float z[3] = {0, 0, 0};
    return h->v.rebuild_topology(n, radius, orig ? orig : z);
}
int gvdb_accumulate_topology(gvdb_handle h, int n, float radius, const float orig[3], int min_depth) {
    H_OR_FAIL(h); NOT_RECORDING(h, "AccumulateTopology");
    const float z[3] = {0, 0, 0};
    return h->v.accumulate_topology(n, radius, orig ? orig : z, min_depth);
}
int gvdb_rebuild_topology_cpu(gvdb_handle h, int n, const float orig[3], const float* host_xyz) {
    H_OR_FAIL(h); NOT_RECORDING(h, "RebuildTopologyCPU");
    const float z[3] = {0, 0, 0};
    return h->v.rebuild_topology_cpu(n, orig ? orig : z, host_xyz);
}
int gvdb_activate_space(gvdb_handle h, const float pos[3]) {
    H_OR_FAIL(h); NOT_RECORDING(h, "ActivateSpace");
    if (!pos) return GVDB_ERR_INVALID;
    return h->v.activate_space(pos);
}
int gvdb_clear(gvdb_handle h) { H_OR_FAIL(h); NOT_RECORDING(h, "Clear"); return h->v.clear(); }
int gvdb_finish_topology(gvdb_handle h, int commit_pool, int compute_bounds) {
    H_OR_FAIL(h); NOT_RECORDING(h, "FinishTopology");
    return h->v.finish_topology(commit_pool != 0, compute_bounds != 0);
}
int gvdb_update_atlas(gvdb_handle h) { H_OR_FAIL(h); NOT_RECORDING(h, "UpdateAtlas"); return h->v.update_atlas(); }
int gvdb_clear_channel(gvdb_handle h, int chan) { H_OR_FAIL(h); return h->v.clear_channel(chan); }
int gvdb_insert_points_subcell(gvdb_handle h, int subcell_size, int n, float radius, const float trans[3], int* out_len) {
    H_OR_FAIL(h);
    if (!out_len) return GVDB_ERR_INVALID;
    const float z[3] = {0, 0, 0};
    return h->v.insert_points_subcell(subcell_size, n, radius, trans ? trans : z, out_len);
}
int gvdb_gather_density(gvdb_handle h, int subcell_size, int n, float radius, const float trans[3], int sc_len,
                        int chan_density, int chan_clr, int accumulate) {
    H_OR_FAIL(h);
    const float z[3] = {0, 0, 0};
    return h->v.gather_density(subcell_size, n, radius, trans ? trans : z, sc_len, chan_density, chan_clr, accumulate != 0);
}
int gvdb_compute(gvdb_handle h, int effect, int chan, int iterations, const float params[3], int update_apron,
                 int skip_aprons, float boundval) {
    H_OR_FAIL(h); NOT_RECORDING(h, "Compute");
    const float z[3] = {0, 0, 0};
    return h->v.compute(effect, chan, iterations, params ? params : z, update_apron != 0, skip_aprons != 0, boundval);
}
int gvdb_compute_kernel(gvdb_handle h, void* fn, int chan, int update_apron, int skip_aprons) {
    H_OR_FAIL(h); NOT_RECORDING(h, "ComputeKernel");
    return h->v.compute_kernel(fn, chan, update_apron != 0, skip_aprons != 0);
}
int gvdb_render_kernel(gvdb_handle h, void* fn, int chan, int rbuf) {
    H_OR_FAIL(h); NOT_RECORDING(h, "RenderKernel");
    return h->v.render_kernel(fn, chan, rbuf);
}
int gvdb_module_load(const char* path, void** mod) {
    if (!path || !mod) return GVDB_ERR_INVALID;
    hipModule_t m = nullptr;
    const hipError_t e = hipModuleLoad(&m, path);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e == hipErrorFileNotFound ? GVDB_ERR_IO : GVDB_ERR_HIP;
    }
    *mod = reinterpret_cast<void*>(m);
    return GVDB_OK;
}
int gvdb_module_function(void* mod, const char* name, void** fn) {
    if (!mod || !name || !fn) return GVDB_ERR_INVALID;
    hipFunction_t f = nullptr;
    if (hipModuleGetFunction(&f, reinterpret_cast<hipModule_t>(mod), name) != hipSuccess) {
        (void)hipGetLastError();
        return GVDB_ERR_INVALID;
    }
    *fn = reinterpret_cast<void*>(f);
    return GVDB_OK;
}
int gvdb_module_unload(void* mod) {
    if (!mod) return GVDB_ERR_INVALID;
    return hipModuleUnload(reinterpret_cast<hipModule_t>(mod)) == hipSuccess ? GVDB_OK : GVDB_ERR_HIP;
}
int gvdb_save_vbx(gvdb_handle h, const char* path) { H_OR_FAIL(h); NOT_RECORDING(h, "SaveVBX"); return h->v.save_vbx(path); }
int gvdb_load_vbx(gvdb_handle h, const char* path, int force_major, int force_minor) {
    H_OR_FAIL(h); NOT_RECORDING(h, "LoadVBX");
    return h->v.load_vbx(path, force_major, force_minor);
}
int gvdb_set_transform(gvdb_handle h, const float pretrans[3], const float scale[3], const float angs[3],
                       const float trans[3]) {
    H_OR_FAIL(h);
    if (!pretrans || !scale || !angs || !trans) return GVDB_ERR_INVALID;
    return h->v.set_transform(pretrans, scale, angs, trans);
}
int gvdb_get_transform(gvdb_handle h, float out12[12]) {
    H_OR_FAIL(h);
    if (!out12) return GVDB_ERR_INVALID;
    h->v.get_transform(out12);
    return GVDB_OK;
}
int gvdb_insert_points_subcell_fp16(gvdb_handle h, int subcell_size, int n, float radius, const float trans[3],
                                    int* out_len) {
    H_OR_FAIL(h); NOT_RECORDING(h, "InsertPointsSubcell_FP16");
    if (!out_len) return GVDB_ERR_INVALID;
    const float z[3] = {0, 0, 0};
    return h->v.insert_points_subcell_fp16(subcell_size, n, radius, trans ? trans : z, out_len);
}
int gvdb_gather_level_set_fp16(gvdb_handle h, int subcell_size, int n, float radius, const float trans[3], int sc_len,
                               int chan_levelset, int chan_clr) {
    H_OR_FAIL(h); NOT_RECORDING(h, "GatherLevelSet_FP16");
    const float z[3] = {0, 0, 0};
    return h->v.gather_level_set_fp16(subcell_size, n, radius, trans ? trans : z, sc_len, chan_levelset, chan_clr);
}
int gvdb_gather_level_set(gvdb_handle h, int subcell_size, int n, float radius, const float trans[3], int sc_len,
                          int chan_levelset, int chan_clr, int accumulate) {
    H_OR_FAIL(h);
    const float z[3] = {0, 0, 0};
    return h->v.gather_level_set(subcell_size, n, radius, trans ? trans : z, sc_len, chan_levelset, chan_clr,
                                 accumulate != 0);
}
int gvdb_update_apron(gvdb_handle h, int chan, float boundval) { H_OR_FAIL(h); return h->v.update_apron(chan, boundval); }
int gvdb_update_apron_noalloc(gvdb_handle h, int chan, float boundval) {
    H_OR_FAIL(h);
    return h->v.update_apron_noalloc(chan, boundval);
}

int gvdb_set_camera_orbit(gvdb_handle h, const float angs[3], const float to[3], float dist, float dolly, float fov,
                          float znear, float zfar) {
    H_OR_FAIL(h);
    if (!angs || !to || !(znear > 0) || !(zfar > znear) || !(fov > 0 && fov < 180)) return GVDB_ERR_INVALID;
    for (int a = 0; a < 3; a++) { h->v.cam.angs[a] = angs[a]; h->v.cam.to[a] = to[a]; }
    h->v.cam.dist = dist; h->v.cam.dolly = dolly; h->v.cam.fov = fov; h->v.cam.znear = znear; h->v.cam.zfar = zfar;
    return GVDB_OK;
}
int gvdb_set_light(gvdb_handle h, const float pos[3]) {
    H_OR_FAIL(h);
    if (!pos) return GVDB_ERR_INVALID;
    for (int a = 0; a < 3; a++) h->v.light_[a] = pos[a];
    return GVDB_OK;
}
int gvdb_set_background(gvdb_handle h, const float rgba[4]) {
    H_OR_FAIL(h);
    if (!rgba) return GVDB_ERR_INVALID;
    for (int a = 0; a < 4; a++) h->v.back_[a] = rgba[a];
    return GVDB_OK;
}
static int set3(float* d, float a, float b, float c) { d[0] = a; d[1] = b; d[2] = c; return GVDB_OK; }
int gvdb_set_extinct(gvdb_handle h, float a, float b, float c) { H_OR_FAIL(h); return set3(h->v.extinct_, a, b, c); }
int gvdb_set_steps(gvdb_handle h, float a, float b, float c) {
    H_OR_FAIL(h);
    if (!(a > 0)) return GVDB_ERR_INVALID;
    return set3(h->v.steps_, a, b, c);
}
int gvdb_set_cutoff(gvdb_handle h, float a, float b, float c) { H_OR_FAIL(h); return set3(h->v.cutoff_, a, b, c); }
int gvdb_set_volume_range(gvdb_handle h, float a, float b, float c) { H_OR_FAIL(h); return set3(h->v.vrange_, a, b, c); }
int gvdb_set_shadow_params(gvdb_handle h, float a, float b, float c) { H_OR_FAIL(h); return set3(h->v.shadow_, a, b, c); }
int gvdb_linear_transfer_func(gvdb_handle h, float t0, float t1, const float a[4], const float b[4]) {
    H_OR_FAIL(h);
    if (!a || !b || !(t1 > t0)) return GVDB_ERR_INVALID;
    return h->v.linear_tf(t0, t1, a, b);
}
int gvdb_commit_transfer_func(gvdb_handle h) { H_OR_FAIL(h); NOT_RECORDING(h, "CommitTransferFunc"); return h->v.commit_tf(); }
int gvdb_add_render_buf(gvdb_handle h, int buf, int w, int hgt, int bpp) { H_OR_FAIL(h); NOT_RECORDING(h, "AddRenderBuf"); return h->v.add_render_buf(buf, w, hgt, bpp); }
int gvdb_render(gvdb_handle h, int shading, int chan, int buf) { H_OR_FAIL(h); return h->v.render(shading, chan, buf); }
int gvdb_read_render_buf(gvdb_handle h, int buf, void* dst) { H_OR_FAIL(h); NOT_RECORDING(h, "ReadRenderBuf"); return h->v.read_render_buf(buf, dst); }
int gvdb_set_render_unroll(int shading, int unroll) {
    if (shading != GVDB_SHADE_VOLUME && shading != GVDB_SHADE_TRILINEAR) return GVDB_ERR_INVALID;
    return gvdb::set_knob(shading == GVDB_SHADE_VOLUME ? gvdb::KNOB_RENDER_UNROLL_EA : gvdb::KNOB_RENDER_UNROLL_ISO,
                          unroll)
               ? GVDB_OK
               : GVDB_ERR_INVALID;
}
int gvdb_tune_render(gvdb_handle h, int shading, int chan, int buf, int* best, float* ms3) {
    H_OR_FAIL(h); NOT_RECORDING(h, "tuning");
    return gvdb::tune_render_unroll(h->v, shading, chan, buf, best, ms3);
}
int gvdb_tune_render_shape(gvdb_handle h, int shading, int chan, int buf, int* best2, float* ms12) {
    H_OR_FAIL(h); NOT_RECORDING(h, "tuning");
    return gvdb::tune_render(h->v, shading, chan, buf, best2, ms12);
}
int gvdb_tune_gather(gvdb_handle h, int subcell_size, int n, float radius, const float trans[3], int sc_len, int chan,
                     int* best_threads, float* ms2) {
    H_OR_FAIL(h); NOT_RECORDING(h, "tuning");
    const float z[3] = {0, 0, 0};
    return gvdb::tune_gather(h->v, subcell_size, n, radius, trans ? trans : z, sc_len, chan, best_threads, ms2);
}
int gvdb_set_tuning(int knob, int value) { return gvdb::set_knob(knob, value) ? GVDB_OK : GVDB_ERR_INVALID; }
int gvdb_get_tuning(int knob, int* value, int* choices8, int* nchoices) {
    if (knob < 0 || knob >= gvdb::KNOB_COUNT) return GVDB_ERR_INVALID;
    if (value) *value = gvdb::knob(knob);
    int c[gvdb::KNOB_MAX_VALUES];
    const int n = gvdb::knob_values(knob, c);
    if (choices8)
        for (int i = 0; i < n; i++) choices8[i] = c[i];
    if (nchoices) *nchoices = n;
    return GVDB_OK;
}
int gvdb_get_scene_block(gvdb_handle h, int w, int hgt, float* out68) {
    H_OR_FAIL(h);
    if (!out68 || w <= 0 || hgt <= 0) return GVDB_ERR_INVALID;
    gvdb::ScnBlock s;
    h->v.scene_block(w, hgt, s);
    memcpy(out68, &s, sizeof(s));
    return GVDB_OK;
}
int gvdb_get_pool(gvdb_handle h, int grp, int lev, void* dst, uint64_t* count, uint64_t* stride) {
    H_OR_FAIL(h); NOT_RECORDING(h, "getNode pools");
    gvdb::u64 c = 0, s = 0;
    int rc = h->v.get_pool(grp, lev, dst, &c, &s);
    if (count) *count = c;
    if (stride) *stride = s;
    return rc;
}
int gvdb_get_info(gvdb_handle h, int* out2, float* outf12) { H_OR_FAIL(h); NOT_RECORDING(h, "reading the tree descriptor"); return h->v.get_info(out2, outf12); }
int gvdb_get_atlas_info(gvdb_handle h, int chan, int* out7) {
    H_OR_FAIL(h);
    if (!out7) return GVDB_ERR_INVALID;
    return h->v.get_atlas_info(chan, out7);
}
int gvdb_read_atlas(gvdb_handle h, int chan, void* dst) { H_OR_FAIL(h); NOT_RECORDING(h, "reading the atlas"); return h->v.read_atlas(chan, dst); }
int gvdb_read_atlas_map(gvdb_handle h, void* dst) { H_OR_FAIL(h); NOT_RECORDING(h, "reading the atlas map"); return h->v.read_atlas_map(dst); }
int gvdb_read_subcells(gvdb_handle h, int which, void* dst, uint64_t* count) {
    H_OR_FAIL(h); NOT_RECORDING(h, "reading subcells");
    gvdb::u64 c = 0;
    int rc = h->v.read_subcells(which, dst, &c);
    if (count) *count = c;
    return rc;
}
int gvdb_get_stream(gvdb_handle h, void** stream) {
    H_OR_FAIL(h);
    if (!stream) return GVDB_ERR_INVALID;
    *stream = h->v.stream;
    return GVDB_OK;
}
int gvdb_synchronize(gvdb_handle h) {
    H_OR_FAIL(h); NOT_RECORDING(h, "synchronize");
    return hipStreamSynchronize(h->v.stream) == hipSuccess ? GVDB_OK : GVDB_ERR_HIP;
}
int gvdb_stage_ms(gvdb_handle h, float* out8) {
    H_OR_FAIL(h); NOT_RECORDING(h, "stage timing");
    if (!out8) return GVDB_ERR_INVALID;
    return h->v.stage_ms(out8);
}
int gvdb_kernel_ms(gvdb_handle h, float* out4) {
    H_OR_FAIL(h); NOT_RECORDING(h, "kernel timing");
    if (!out4) return GVDB_ERR_INVALID;
    return h->v.kernel_ms(out4);
}

// ------------------------------------------------------------- batches
int gvdb_batch_create(gvdb_batch* out, int device, int max_grids) {
    if (!out) return GVDB_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return GVDB_ERR_HIP;
    gvdb_batch_s* b = new (std::nothrow) gvdb_batch_s(device, max_grids);
    if (!b) return GVDB_ERR_OOM;
    int rc = b->b.init();
    if (rc) { delete b; return rc; }
    *out = b;
    return GVDB_OK;
}
int gvdb_batch_destroy(gvdb_batch b) {
    if (!b) return GVDB_ERR_INVALID;
    delete b;
    return GVDB_OK;
}
const char* gvdb_batch_last_error(gvdb_batch b) { return b ? b->b.err.c_str() : "null batch"; }

// handles -> volumes (null handles are rejected)
#define VOLS(b, count, hs)                                                       \
    if (!(b) || !(hs) || (count) < 1) return GVDB_ERR_INVALID;                   \
    std::vector<gvdb::Volume*> vs_(static_cast<size_t>(count));                             \
    for (int i_ = 0; i_ < (count); i_++) {                                       \
        if (!(hs)[i_]) return GVDB_ERR_INVALID;                                  \
        vs_[i_] = &(hs)[i_]->v;                                                  \
    }

int gvdb_batch_rebuild_topology(gvdb_batch b, int count, const gvdb_handle* hs, const int* n, const float* radius,
                                const float* orig3, int* failed) {
    VOLS(b, count, hs);
    if (!n || !radius) return GVDB_ERR_INVALID;
    std::vector<float> z(size_t(count) * 3, 0.f);
    return b->b.rebuild_topology(vs_.data(), count, n, radius, orig3 ? orig3 : z.data(), failed);
}
int gvdb_batch_finish_topology(gvdb_batch b, int count, const gvdb_handle* hs, int commit_pool, int compute_bounds,
                               int* failed) {
    VOLS(b, count, hs);
    return b->b.finish_topology(vs_.data(), count, commit_pool != 0, compute_bounds != 0, failed);
}
int gvdb_batch_update_atlas(gvdb_batch b, int count, const gvdb_handle* hs, int* failed) {
    VOLS(b, count, hs);
    return b->b.update_atlas(vs_.data(), count, failed);
}
int gvdb_batch_clear_channel(gvdb_batch b, int count, const gvdb_handle* hs, int chan, int* failed) {
    VOLS(b, count, hs);
    return b->b.clear_channel(vs_.data(), count, chan, failed);
}
int gvdb_batch_insert_points_subcell(gvdb_batch b, int count, const gvdb_handle* hs, int subcell_size, const int* n,
                                     const float* radius, const float* trans3, int* out_len, int* failed) {
    VOLS(b, count, hs);
    if (!n || !radius || !out_len) return GVDB_ERR_INVALID;
    std::vector<float> z(size_t(count) * 3, 0.f);
    return b->b.insert_points_subcell(vs_.data(), count, subcell_size, n, radius, trans3 ? trans3 : z.data(), out_len,
                                      failed);
}
int gvdb_batch_gather_density(gvdb_batch b, int count, const gvdb_handle* hs, int subcell_size, const int* sc_len,
                              const float* radius, int chan_density, int accumulate, int* failed) {
    VOLS(b, count, hs);
    if (!sc_len || !radius) return GVDB_ERR_INVALID;
    return b->b.gather_density(vs_.data(), count, subcell_size, sc_len, radius, chan_density, accumulate != 0, failed);
}
int gvdb_batch_gather_level_set(gvdb_batch b, int count, const gvdb_handle* hs, int subcell_size, const int* sc_len,
                                const float* radius, int chan_levelset, int accumulate, int* failed) {
    VOLS(b, count, hs);
    if (!sc_len || !radius) return GVDB_ERR_INVALID;
    return b->b.gather_level_set(vs_.data(), count, subcell_size, sc_len, radius, chan_levelset, accumulate != 0, failed);
}
int gvdb_batch_update_apron(gvdb_batch b, int count, const gvdb_handle* hs, int chan, float boundval, int* failed) {
    VOLS(b, count, hs);
    return b->b.update_apron(vs_.data(), count, chan, boundval, failed);
}
int gvdb_batch_render(gvdb_batch b, int count, const gvdb_handle* hs, const int* shading, const int* chan,
                      const int* buf, int* failed) {
    VOLS(b, count, hs);
    if (!shading || !chan || !buf) return GVDB_ERR_INVALID;
    return b->b.render(vs_.data(), count, shading, chan, buf, failed);
}
int gvdb_batch_tune(gvdb_batch b, int count, const gvdb_handle* hs, int stage, int subcell_size, const int* sc_len,
                    const float* radius, int chan, float boundval, const int* shading, const int* chans,
                    const int* bufs, int* best, float* ms, int* failed) {
    VOLS(b, count, hs);
    if (stage == 0 && (!sc_len || !radius)) return GVDB_ERR_INVALID;
    if (stage == 2 && (!shading || !chans || !bufs)) return GVDB_ERR_INVALID;
    return gvdb::tune_batch(b->b, vs_.data(), count, stage, subcell_size, sc_len, radius, chan, boundval, shading,
                            chans, bufs, best, ms, failed);
}
int gvdb_batch_synchronize(gvdb_batch b) {
    if (!b) return GVDB_ERR_INVALID;
    return b->b.synchronize();
}
int gvdb_batch_elapsed_ms(gvdb_batch b, float* ms) {
    if (!b || !ms) return GVDB_ERR_INVALID;
    return b->b.elapsed_ms(ms);
}
int gvdb_release_stream(gvdb_handle h) {
    H_OR_FAIL(h); NOT_RECORDING(h, "releasing the stream");
    return h->v.use_stream(std::shared_ptr<gvdb::SharedStream>());
}

// ------------------------------------------------------------- serving
int gvdb_server_create(gvdb_server* out, int device, unsigned long long hbm_budget, int max_batch) {
    if (!out) return GVDB_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return GVDB_ERR_HIP;
    gvdb_server_s* s = new (std::nothrow) gvdb_server_s(device, hbm_budget, max_batch);
    if (!s) return GVDB_ERR_OOM;
    int rc = s->s.init();
    if (rc) { delete s; return rc; }
    *out = s;
    return GVDB_OK;
}
int gvdb_server_destroy(gvdb_server s) {
    if (!s) return GVDB_ERR_INVALID;
    delete s;
    return GVDB_OK;
}
const char* gvdb_server_last_error(gvdb_server s) { return s ? s->s.err.c_str() : "null server"; }
int gvdb_server_add_grid(gvdb_server s, gvdb_handle h, const float* host_xyz, int n, float radius) {
    if (!s || !h) return GVDB_ERR_INVALID;
    return s->s.add_grid(&h->v, host_xyz, n, radius);
}
int gvdb_server_reset(gvdb_server s) {
    if (!s) return GVDB_ERR_INVALID;
    return s->s.reset();
}
int gvdb_server_submit_render(gvdb_server s, int grid, double arrival, int w, int hgt, int shading, const float angs[3],
                              float dist, unsigned char* out_rgba) {
    if (!s) return GVDB_ERR_INVALID;
    return s->s.submit_render(grid, arrival, w, hgt, shading, angs, dist, out_rgba);
}
int gvdb_server_submit_insert(gvdb_server s, int grid, double arrival, const float* host_xyz, int n) {
    if (!s) return GVDB_ERR_INVALID;
    return s->s.submit_insert(grid, arrival, host_xyz, n);
}
int gvdb_server_clear_requests(gvdb_server s) {
    if (!s) return GVDB_ERR_INVALID;
    s->s.clear_requests();
    return GVDB_OK;
}
int gvdb_server_run(gvdb_server s, int mode, double* latency, int* status, double* wall) {
    if (!s || (mode != 0 && mode != 1)) return GVDB_ERR_INVALID;
    return s->s.run(mode, latency, status, wall);
}
unsigned long long gvdb_server_device_bytes(gvdb_server s) { return s ? s->s.device_bytes() : 0ull; }

// ------------------------------------------------------------- graphs
int gvdb_capture_begin(gvdb_handle h, gvdb_graph* out) {
    if (!out) return GVDB_ERR_INVALID;
    *out = nullptr;
    H_OR_FAIL(h);
    gvdb_graph_s* g = new (std::nothrow) gvdb_graph_s();
    if (!g) return GVDB_ERR_OOM;
    gvdb::Graph* raw = nullptr;
    int rc = gvdb::capture_begin(h->v, raw, h->v.err);
    if (rc) { delete g; return rc; }
    g->g.reset(raw);
    *out = g;
    return GVDB_OK;
}
int gvdb_capture_end(gvdb_handle h, gvdb_graph g) {
    H_OR_FAIL(h);
    if (!g || !g->g) return GVDB_ERR_INVALID;
    int rc = gvdb::capture_end(h->v, g->g.get(), g->err);
    if (rc) h->v.err = g->err;
    return rc;
}
int gvdb_batch_capture_begin(gvdb_batch b, gvdb_graph* out) {
    if (!b || !out) return GVDB_ERR_INVALID;
    *out = nullptr;
    gvdb_graph_s* g = new (std::nothrow) gvdb_graph_s();
    if (!g) return GVDB_ERR_OOM;
    gvdb::Graph* raw = nullptr;
    int rc = b->b.capture_begin(raw);
    if (rc) { delete g; return rc; }
    g->g.reset(raw);
    *out = g;
    return GVDB_OK;
}
int gvdb_batch_capture_end(gvdb_batch b, gvdb_graph g) {
    if (!b || !g || !g->g) return GVDB_ERR_INVALID;
    int rc = b->b.capture_end(g->g.get());
    if (rc) g->err = b->b.err;
    return rc;
}
int gvdb_graph_launch(gvdb_graph g) {
    if (!g || !g->g) return GVDB_ERR_INVALID;
    return gvdb::graph_launch(g->g.get(), g->err);
}
int gvdb_graph_synchronize(gvdb_graph g) {
    if (!g || !g->g) return GVDB_ERR_INVALID;
    return gvdb::graph_synchronize(g->g.get(), g->err);
}
int gvdb_graph_destroy(gvdb_graph g) {
    if (!g) return GVDB_ERR_INVALID;
    delete g;
    return GVDB_OK;
}
const char* gvdb_graph_last_error(gvdb_graph g) { return g ? g->err.c_str() : "null graph"; }

}  // extern "C"
