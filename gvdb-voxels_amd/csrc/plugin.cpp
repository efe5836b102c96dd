// plugin.cpp -- user kernels over the volume: ComputeKernel and RenderKernel
// (src/gvdb_volume_gvdb.cpp:4499-4525 and :4309-4333 of the reference).
//
// The reference hands user kernels its VDBInfo block and CUDA surfaces.  A
// kernel here gets a gvdb_device_view (include/gvdb_amd_device.h): the
// channel's brick-major atlas with its geometry, the atlas map, the node
// pools and, for render kernels, the scene block.  The launch shapes are the
// reference's.
#include "volume.h"
#include "../../include/gvdb_amd.h"
#include "../../include/gvdb_amd_device.h"
#include <hip/hip_runtime.h>
#include <cstring>

namespace gvdb {

#define PLG_HIPCHK(x)                                      \
    do {                                                   \
        hipError_t e_ = (x);                               \
        if (e_ != hipSuccess) return hip_fail(e_, #x);     \
    } while (0)

static_assert(MAXLEV == GVDB_VIEW_MAXLEV, "device view levels");

int Volume::upload_view(int c, int rbuf, void** dev) {
    gvdb_device_view hv;
    memset(&hv, 0, sizeof(hv));
    int rb = resolve_bounds();
    if (rb) return rb;
    const Tree t = make_tree();
    const Channel& ch = chan_[c];
    hv.atlas = ch.data.p;
    hv.type = ch.type;
    hv.brick_res = t.brick_res;
    hv.brickwid = t.brickwid;
    hv.apron = t.apron;
    for (int a = 0; a < 3; a++) {
        hv.atlas_cnt[a] = atlas_cnt_[a];
        hv.atlas_res[a] = atlas_cnt_[a] * t.brick_res;
    }
    hv.used = int(atlas_used_);
    hv.amap = amap_.as<int32_t>();
    hv.levels = levels_;
    hv.top_lev = top_lev_;
    for (int l = 0; l < levels_; l++) {
        hv.res[l] = t.res[l];
        hv.range[l] = t.range[l];
        hv.nodecnt[l] = t.nodecnt[l];
        hv.nodes[l] = t.nodes[l];
        hv.clists[l] = reinterpret_cast<const uint64_t*>(t.clists[l]);
    }
    if (rbuf >= 0) {
        ScnBlock sb;
        scene_block(rbuf_[rbuf].w, rbuf_[rbuf].h, sb);
        PLG_HIPCHK(view_scene_.ensure(sizeof(sb)));
        PLG_HIPCHK(hipMemcpyAsync(view_scene_.p, &sb, sizeof(sb), hipMemcpyHostToDevice, stream));
        hv.scene = view_scene_.as<float>();
        hv.width = rbuf_[rbuf].w;
        hv.height = rbuf_[rbuf].h;
    }
    PLG_HIPCHK(view_dev_.ensure(sizeof(hv)));
    PLG_HIPCHK(hipMemcpyAsync(view_dev_.p, &hv, sizeof(hv), hipMemcpyHostToDevice, stream));
    PLG_HIPCHK(hipStreamSynchronize(stream));   // the host copies above live on this frame
    *dev = view_dev_.p;
    return GVDB_OK;
}

int Volume::compute_kernel(void* fn, int c, bool update_apron_after, bool skip_aprons) {
    if (!fn) return fail(GVDB_ERR_INVALID, "ComputeKernel: null kernel");
    if (c < 0 || c >= 8 || !chan_[c].exists) return fail(GVDB_ERR_INVALID, "ComputeKernel: no such channel");
    int rc = commit_pools();
    if (rc) return rc;
    void* view = nullptr;
    rc = upload_view(c, -1, &view);
    if (rc) return rc;
    const int br = brickwid_ + 2 * apron_;
    int3 res = make_int3(atlas_cnt_[0] * br, atlas_cnt_[1] * br, atlas_cnt_[2] * br);
    const int cnt[3] = {skip_aprons ? atlas_cnt_[0] * brickwid_ : res.x, skip_aprons ? atlas_cnt_[1] * brickwid_ : res.y,
                        skip_aprons ? atlas_cnt_[2] * brickwid_ : res.z};
    unsigned char ch = (unsigned char)c;
    void* args[3] = {&view, &res, &ch};
    PLG_HIPCHK(hipModuleLaunchKernel(reinterpret_cast<hipFunction_t>(fn), unsigned((cnt[0] + 7) / 8),
                                     unsigned((cnt[1] + 7) / 8), unsigned((cnt[2] + 7) / 8), 8, 8, 8, 0, stream, args,
                                     nullptr));
    atlas_epoch_++;
    if (update_apron_after) {
        // UpdateApron() with no arguments: every channel, boundary value 0
        for (int k = 0; k < 8; k++) {
            if (!chan_[k].exists) continue;
            rc = update_apron(k, 0.0f);
            if (rc) return rc;
        }
    }
    return GVDB_OK;
}

int Volume::render_kernel(void* fn, int c, int b) {
    if (!fn) return fail(GVDB_ERR_INVALID, "RenderKernel: null kernel");
    if (c < 0 || c >= 8 || !chan_[c].exists) return fail(GVDB_ERR_INVALID, "RenderKernel: no such channel");
    if (b < 0 || b >= 4 || rbuf_[b].w == 0) return fail(GVDB_ERR_INVALID, "RenderKernel: no such render buffer");
    int rc = commit_pools();
    if (rc) return rc;
    void* view = nullptr;
    rc = upload_view(c, b, &view);
    if (rc) return rc;
    unsigned char ch = (unsigned char)c;
    void* out = rbuf_[b].data.p;
    void* args[3] = {&view, &ch, &out};
    PLG_HIPCHK(hipModuleLaunchKernel(reinterpret_cast<hipFunction_t>(fn), unsigned((rbuf_[b].w + 7) / 8),
                                     unsigned((rbuf_[b].h + 7) / 8), 1, 8, 8, 1, 0, stream, args, nullptr));
    return GVDB_OK;
}

}  // namespace gvdb
