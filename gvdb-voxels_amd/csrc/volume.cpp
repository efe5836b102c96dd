// volume.cpp -- host runtime of one grid and the C ABI (include/gvdb_amd.h).
//
// Follows the host side of VolumeGVDB (source/gvdb_library/src/
// gvdb_volume_gvdb.cpp of the reference) for the point-to-volume path:
// Configure, pools with doubling growth, GPU RebuildTopology, the host
// ActivateSpace build, FinishTopology, UpdateAtlas, InsertPointsSubcell,
// GatherDensity, UpdateApron and Render.  All device work runs on the grid's
// own stream; the only host round trips are the ones whose results the API
// returns to the caller (bounding box/depth, node counts, subcell entries).
#include "volume.h"
#include "graph.h"
#include "../../include/gvdb_amd.h"
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <algorithm>

namespace gvdb {

// ------------------------------------------------------------- buffers ----
hipError_t DevBuf::ensure(size_t n, bool preserve, hipStream_t s) {
    if (n <= bytes) return hipSuccess;
    bump_alloc_generation();
    size_t want = std::max(n, bytes + bytes / 4);
    want = (want + 255) & ~size_t(255);
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        want = (n + 255) & ~size_t(255);
        e = hipMalloc(&q, want);
        if (e != hipSuccess) return e;
    }
    if (preserve && p && bytes) {
        e = hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { (void)hipFree(q); return e; }
    }
    if (p) (void)hipFree(p);
    p = q;
    bytes = want;
    return hipSuccess;
}

void DevBuf::release() {
    if (p) bump_alloc_generation();
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

#define HIPCHK(x)                                          \
    do {                                                   \
        hipError_t e_ = (x);                               \
        if (e_ != hipSuccess) return hip_fail(e_, #x);     \
    } while (0)

Exec Volume::self_exec() {
    if (capture_) return Exec{stream, &capture_->arena, small_.as<uint32_t>(), pinned_, capture_};
    return Exec{stream, &arena_, small_.as<uint32_t>(), pinned_};
}

int Volume::finish_recorded_insert(const uint32_t* slot) {
    // the words the insert's drain copies: linked bricks, points outside
    // every brick, the 64-bit entry total
    if (slot[SM_ORPHANS] != 0)
        return fail(GVDB_ERR_STATE, "graph replay: " + std::to_string(slot[SM_ORPHANS]) +
                                        " points fell outside the topology, which the recorded insert leaves out; "
                                        "rebuild the topology (or insert without the graph)");
    u64 total = 0;
    memcpy(&total, slot + SM_LEN, 8);
    if (total > 0x7FFFFFFFull) return fail(GVDB_ERR_RANGE, "graph replay: more than 2^31 subcell entries");
    sc_len_ = int(total);
    sc_bricks_ = int(slot[SM_LINKED]);
    return GVDB_OK;
}

int Volume::not_recordable(const char* what) {
    err = std::string(what) + ": cannot be recorded into a graph (the per-frame chain is ClearChannel, "
                              "InsertPointsSubcell, GatherDensity/GatherLevelSet, UpdateApron and Render)";
    if (capture_ && !capture_->failed) {
        capture_->failed = true;
        capture_->err = err;
    }
    return GVDB_ERR_STATE;
}

void Volume::forget_graph(Graph* g) {
    graphs_.erase(std::remove(graphs_.begin(), graphs_.end(), g), graphs_.end());
}

SharedStream::~SharedStream() {
    if (!s) return;
    (void)hipSetDevice(device);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
}

std::shared_ptr<SharedStream> SharedStream::make(int device, hipError_t& e) {
    auto p = std::make_shared<SharedStream>();
    p->device = device;
    e = hipStreamCreateWithFlags(&p->s, hipStreamNonBlocking);
    if (e != hipSuccess) p->s = nullptr;
    return p;
}

int Volume::use_stream(const std::shared_ptr<SharedStream>& s) {
    const std::shared_ptr<SharedStream>& to = s ? s : own_stream_;
    if (to->s == stream) return GVDB_OK;
    HIPCHK(hipEventRecord(handoff_, stream));
    HIPCHK(hipStreamWaitEvent(to->s, handoff_, 0));
    cur_stream_ = to;
    stream = to->s;
    return GVDB_OK;
}

int Volume::hip_fail(hipError_t e, const char* what) {
    err = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
    return e == hipErrorOutOfMemory ? GVDB_ERR_OOM : GVDB_ERR_HIP;
}

Volume::Volume(int dev) : device(dev) {}

Volume::~Volume() {
    // graphs that replay calls on this grid can no longer run
    for (Graph* g : graphs_) {
        g->failed = true;
        g->err = "a grid it replays calls on was destroyed";
        for (size_t k = g->vols.size(); k-- > 0;)
            if (g->vols[k] == this) {
                g->vols.erase(g->vols.begin() + long(k));
                g->topo.erase(g->topo.begin() + long(k));
            }
        g->inserts.erase(std::remove_if(g->inserts.begin(), g->inserts.end(),
                                        [this](const Graph::Insert& in) { return in.v == this; }),
                         g->inserts.end());
    }
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    for (int g = 0; g < 2; g++)
        for (int l = 0; l < MAXLEV; l++) pool_[g][l].dev.release();
    for (auto& c : chan_) c.data.release();
    for (auto& r : rbuf_) r.data.release();
    DevBuf* bufs[] = {&amap_, &nbr_, &nmask_, &nbase_, &faces_, &lenparts_, &pts_owned_, &clr_owned_, &sc_cnt_, &sc_off_, &sc_pos_, &sc_nid_, &sc_map_, &sc_pnt_, &sc_use_,
                      &bkt_key_, &bkt_slot_, &bkt_cnt_, &bkt_off_, &sorted_, &sc_flags_, &small_, &masks_, &bitmaps_, &ranks_, &scan_tmp_, &flags_, &slot_scan_, &tf_dev,
                      &incr_marker_, &sc_pnt16_, &view_dev_, &view_scene_, &kb_keys_, &kb_keysB_, &kb_cnt_, &kb_off_,
                      &kb_tmp_, &kb_hist_, &kb_flags_, &kb_pos_, &kb_list_, &btab_, &bbox_part_};
    for (DevBuf* b : bufs) b->release();
    for (int s = 0; s < ST_COUNT; s++)
        for (int k = 0; k < 2; k++)
            if (ev_[s][k]) (void)hipEventDestroy(ev_[s][k]);
    for (int s = 0; s < 4; s++)
        for (int k = 0; k < 2; k++)
            if (kev_[s][k]) (void)hipEventDestroy(kev_[s][k]);
    if (bounds_ev_) (void)hipEventDestroy(bounds_ev_);
    if (small_ev_) (void)hipEventDestroy(small_ev_);
    if (pinned_) (void)hipHostFree(pinned_);
    arena_.release();
    if (handoff_) (void)hipEventDestroy(handoff_);
    cur_stream_.reset();
    own_stream_.reset();
}

int Volume::init() {
    HIPCHK(hipSetDevice(device));
    {
        hipError_t e = hipSuccess;
        own_stream_ = SharedStream::make(device, e);
        HIPCHK(e);
    }
    cur_stream_ = own_stream_;
    stream = own_stream_->s;
    for (int s = 0; s < ST_COUNT; s++)
        for (int k = 0; k < 2; k++) HIPCHK(hipEventCreate(&ev_[s][k]));
    for (int s = 0; s < 4; s++)
        for (int k = 0; k < 2; k++) HIPCHK(hipEventCreate(&kev_[s][k]));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&pinned_), 256 * sizeof(uint32_t), hipHostMallocDefault));
    HIPCHK(small_.ensure(SMALL_WORDS * 4 * 4));
    HIPCHK(arena_.init(1 << 16));
    HIPCHK(hipEventCreateWithFlags(&handoff_, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&bounds_ev_, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&small_ev_, hipEventDisableTiming));
    // the device's copy of the allocation-free apron's scratch array, found
    // once per device here so that entry point never has to look it up
    HIPCHK(noalloc_scratch(device, &noalloc_));
    // Initialize(): default transfer function and range (gvdb_volume_gvdb.cpp:2341-2348)
    tf_host.assign(size_t(TF_SIZE) * 4, 0.f);
    const float a[4] = {0, 0, 0, 0}, b[4] = {1, 1, 1, 0.1f};
    linear_tf(0.f, 1.f, a, b);
    return GVDB_OK;
}

void Volume::stage_begin(Stage s) {
    if (capture_) return;   // no timing inside a recording
    (void)hipEventRecord(ev_[s][0], stream);
}
void Volume::stage_end(Stage s) {
    if (capture_) return;
    (void)hipEventRecord(ev_[s][1], stream);
    ev_used_[s] = true;
}

int Volume::stage_ms(float* out) {
    HIPCHK(hipStreamSynchronize(stream));
    for (int s = 0; s < ST_COUNT; s++) {
        out[s] = 0.f;
        if (ev_used_[s]) HIPCHK(hipEventElapsedTime(&out[s], ev_[s][0], ev_[s][1]));
    }
    out[ST_COUNT] = 0.f;
    return GVDB_OK;
}

int Volume::kernel_ms(float* out) {
    HIPCHK(hipStreamSynchronize(stream));
    for (int s = 0; s < 4; s++) {
        out[s] = 0.f;
        if (kev_used_[s]) HIPCHK(hipEventElapsedTime(&out[s], kev_[s][0], kev_[s][1]));
    }
    return GVDB_OK;
}

// ------------------------------------------------------- configuration ----
int Volume::configure(int levels, const int* log2dims, const int* initcnt) {
    if (levels < 2 || levels > MAXLEV) return fail(GVDB_ERR_INVALID, "Configure: levels must be 2..10");
    HIPCHK(hipStreamSynchronize(stream));
    levels_ = levels;
    int defcnt[MAXLEV];
    {
        // Configure(q4..q0) pool sizes (gvdb_volume_gvdb.cpp:2369-2374)
        int c = 4;
        defcnt[0] = 4;
        for (int l = 1; l < levels; l++) { defcnt[l] = c; c >>= 1; }
    }
    for (int n = 0; n < levels; n++) {
        int r = log2dims[n];
        if (r < 0 || r > 10) return fail(GVDB_ERR_INVALID, "Configure: log2 dims must be 0..10");
        logdim_[n] = (r == 0) ? 1 : r;
        res_[n] = 1 << logdim_[n];
    }
    for (int l = 0; l < levels; l++) {
        long long rr = res_[0];
        for (int k = 1; k <= l; k++) rr *= res_[k];
        range_[l] = rr;
    }
    if (range_[levels - 1] > (1ll << 30)) return fail(GVDB_ERR_INVALID, "Configure: tree range exceeds 2^30 voxels");
    for (int g = 0; g < 2; g++)
        for (int l = 0; l < MAXLEV; l++) {
            pool_[g][l].dev.release();
            pool_[g][l] = Pool();
        }
    for (int n = 0; n < levels; n++) {
        u64 cnt = u64(std::max(1, initcnt ? initcnt[n] : defcnt[n]));
        Pool& p0 = pool_[0][n];
        p0.exists = true;
        p0.stride = sizeof(Node);
        p0.max = cnt;
        if (n > 0) {
            Pool& p1 = pool_[1][n];
            p1.exists = true;
            p1.stride = sizeof(u64) * voxcnt(n);
            p1.max = cnt;
        }
    }
    brickwid_ = res_[0];
    host_valid_ = true;
    dev_stale_ = false;
    return clear();
}

int Volume::set_channel_default(int x, int y, int z) {
    if (x <= 0 || y <= 0 || z <= 0) return fail(GVDB_ERR_INVALID, "SetChannelDefault: counts must be positive");
    default_axiscnt_[0] = x; default_axiscnt_[1] = y; default_axiscnt_[2] = z;
    return GVDB_OK;
}

int Volume::add_channel(int c, int dtype, int apron, int filter, int border, int ax, int ay, int az) {
    // AddChannel + AtlasCreate (gvdb_volume_gvdb.cpp:2437-2453, gvdb_allocator.cpp:528-560)
    if (levels_ == 0) return fail(GVDB_ERR_STATE, "AddChannel before Configure");
    if (c < 0 || c >= 8) return fail(GVDB_ERR_INVALID, "AddChannel: channel must be 0..7");
    if (dtype != T_FLOAT && dtype != T_HALF && dtype != T_UCHAR4)
        return fail(GVDB_ERR_UNSUPPORTED, "AddChannel: float, half and uchar4 channels");
    if (apron < 0 || apron > 1) return fail(GVDB_ERR_UNSUPPORTED, "AddChannel: apron must be 0 or 1");
    if (c > 0 && !chan_[0].exists) return fail(GVDB_ERR_STATE, "AddChannel: channel 0 first");
    if (ax == 0 && ay == 0 && az == 0) {
        if (c == 0) { ax = default_axiscnt_[0]; ay = default_axiscnt_[1]; az = default_axiscnt_[2]; }
        else { ax = atlas_cnt_[0]; ay = atlas_cnt_[1]; az = atlas_cnt_[2]; }
    }
    if (c == 0) {
        atlas_cnt_[0] = ax; atlas_cnt_[1] = ay; atlas_cnt_[2] = az;
        atlas_max_ = u64(ax) * ay * az;
        apron_ = apron;
    } else if (ax != atlas_cnt_[0] || ay != atlas_cnt_[1] || az != atlas_cnt_[2] || apron != apron_) {
        return fail(GVDB_ERR_UNSUPPORTED, "AddChannel: all channels share channel 0's atlas layout");
    }
    atlas_epoch_++;
    Channel& ch = chan_[c];
    ch.exists = true;
    ch.type = dtype;
    ch.filter = filter;
    ch.border = border;
    int br = brickwid_ + 2 * apron_;
    size_t bytes = size_t(atlas_max_) * br * br * br * ch.elem();
    HIPCHK(ch.data.ensure(bytes));
    HIPCHK(launch_fill_u32(ch.data.as<uint32_t>(), 0u, ch.data.bytes / 4, stream));
    return GVDB_OK;
}

int Volume::set_epsilon(float eps, int maxiter) {
    epsilon_ = eps;
    maxiter_ = maxiter;
    return GVDB_OK;
}

// ---------------------------------------------------------------- points ----
int Volume::set_points_gpu(const void* p, int n, int off, int stride) {
    if (n < 0 || (n > 0 && !p)) return fail(GVDB_ERR_INVALID, "SetPoints: null positions");
    if (stride < 12 || off < 0) return fail(GVDB_ERR_INVALID, "SetPoints: stride must hold a float3");
    pts_ = PointSrc{reinterpret_cast<const char*>(p), n, off, stride};
    return GVDB_OK;
}

int Volume::set_colors_gpu(const void* p, int n, int off, int stride) {
    // SetPoints' colour DataPtr (gvdb_volume_gvdb.cpp:4900-4908): a uint32
    // (uchar4) per point; null clears it
    if (!p || n <= 0) {
        clr_ = PointSrc{nullptr, 0, 0, 4};
        return GVDB_OK;
    }
    if (stride < 4 || off < 0) return fail(GVDB_ERR_INVALID, "SetPoints: colour stride must hold a uint32");
    clr_ = PointSrc{reinterpret_cast<const char*>(p), n, off, stride};
    return GVDB_OK;
}

int Volume::upload_colors(const uint32_t* host, int n) {
    if (!host || n <= 0) return set_colors_gpu(nullptr, 0, 0, 4);
    HIPCHK(clr_owned_.ensure(size_t(n) * 4));
    HIPCHK(hipMemcpyAsync(clr_owned_.p, host, size_t(n) * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    clr_ = PointSrc{reinterpret_cast<const char*>(clr_owned_.p), n, 0, 4};
    return GVDB_OK;
}

int Volume::set_color_channel(int c) {
    // SetColorChannel (gvdb_volume_gvdb.cpp:686-689); 255 (CHAN_UNDEF) or -1: none
    if (c < 0 || c == CHAN_UNDEF) {
        clr_chan_ = CHAN_UNDEF;
        return GVDB_OK;
    }
    if (c >= 8) return fail(GVDB_ERR_INVALID, "SetColorChannel: channel 0..7");
    clr_chan_ = c;
    return GVDB_OK;
}

int Volume::upload_points(const float* host, int n) {
    if (n < 0 || (n > 0 && !host)) return fail(GVDB_ERR_INVALID, "upload_points: null positions");
    HIPCHK(pts_owned_.ensure(size_t(std::max(n, 1)) * 12));
    if (n) {
        HIPCHK(hipMemcpyAsync(pts_owned_.p, host, size_t(n) * 12, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    pts_ = PointSrc{reinterpret_cast<const char*>(pts_owned_.p), n, 0, 12};
    return GVDB_OK;
}

// ----------------------------------------------------------------- pools ----
u64 Volume::pool_alloc(int grp, int lev) {
    // PoolAlloc (gvdb_allocator.cpp:163-192) on the host mirror
    if (lev >= levels_) return ID_UNDEFL;
    Pool& p = pool_[grp][lev];
    if (!p.exists) return ID_UNDEFL;
    if (p.lastEle >= p.max) p.max *= 2;
    if (p.host.size() < p.max * p.stride) p.host.resize(p.max * p.stride, 0);
    p.lastEle++;
    p.usedNum++;
    return Elem(grp, lev, p.lastEle - 1);
}

int Volume::ensure_pool_capacity(int grp, int lev, u64 need, bool preserve, hipStream_t s) {
    Pool& p = pool_[grp][lev];
    while (p.max < need) p.max *= 2;
    HIPCHK(p.dev.ensure(size_t(p.max * p.stride), preserve, s));
    return GVDB_OK;
}

Node* Volume::hnode(u64 ref) {
    Pool& p = pool_[0][(ref >> 8) & 0xFF];
    return reinterpret_cast<Node*>(&p.host[(ref >> 16) * p.stride]);
}
u64* Volume::hclist(u64 ref) {
    Pool& p = pool_[ref & 0xFF][(ref >> 8) & 0xFF];
    return reinterpret_cast<u64*>(&p.host[(ref >> 16) * p.stride]);
}

int Volume::sync_host() {
    // PoolFetchAll (gvdb_allocator.cpp:119-131)
    if (host_valid_) return GVDB_OK;
    HIPCHK(hipStreamSynchronize(stream));
    for (int g = 0; g < 2; g++)
        for (int l = 0; l < levels_; l++) {
            Pool& p = pool_[g][l];
            if (!p.exists) continue;
            if (p.host.size() < p.max * p.stride) p.host.resize(p.max * p.stride, 0);
            if (p.lastEle) HIPCHK(hipMemcpy(p.host.data(), p.dev.p, p.lastEle * p.stride, hipMemcpyDeviceToHost));
        }
    host_valid_ = true;
    return GVDB_OK;
}

int Volume::commit_pools() {
    // PoolCommitAll (gvdb_allocator.cpp:93-116)
    if (!dev_stale_) return GVDB_OK;
    if (capture_) return not_recordable("pools built on the host and not yet on the device");
    for (int g = 0; g < 2; g++)
        for (int l = 0; l < levels_; l++) {
            Pool& p = pool_[g][l];
            if (!p.exists) continue;
            int rc = ensure_pool_capacity(g, l, std::max<u64>(p.lastEle, 1));
            if (rc) return rc;
            if (p.lastEle)
                HIPCHK(hipMemcpyAsync(p.dev.p, p.host.data(), p.lastEle * p.stride, hipMemcpyHostToDevice, stream));
        }
    HIPCHK(hipStreamSynchronize(stream));
    dev_stale_ = false;
    return GVDB_OK;
}

void Volume::setup_node(u64 id, int lev, const int pos[3]) {
    // SetupNode (gvdb_volume_gvdb.cpp:2512-2524)
    Node* n = hnode(id);
    memset(n, 0, sizeof(Node));
    n->mLev = uint8_t(lev);
    n->mPos[0] = pos[0]; n->mPos[1] = pos[1]; n->mPos[2] = pos[2];
    n->mChildList = ID_UNDEFL;
    n->mParent = ID_UNDEFL;
    n->mValue[0] = n->mValue[1] = n->mValue[2] = -1;
    n->mFlags = 1;
}

void Volume::compute_top_lev() {
    // PrepareVDB's top level: the lowest level holding exactly one node (gvdb_volume_gvdb.cpp:3955-3979)
    int t = 1;
    for (int n = levels_ - 1; n >= 0; n--)
        if (pool_[0][n].lastEle == 1) t = n;
    top_lev_ = t;
}

int Volume::clear() {
    // Clear (gvdb_volume_gvdb.cpp:2487-2503)
    for (int g = 0; g < 2; g++)
        for (int l = 0; l < levels_; l++) { pool_[g][l].lastEle = 0; pool_[g][l].usedNum = 0; }
    root_ = ID_UNDEFL;
    atlas_used_ = 0;
    rebuild_topo_ = true;
    host_valid_ = true;
    dev_stale_ = false;
    topo_gen_++;
    return GVDB_OK;
}

void Volume::covering_node_host(int lev, const int pos[3], int out[3]) const {
    // GetCoveringNode (gvdb_volume_gvdb.cpp:2915-2931)
    long long rg = range_[lev];
    for (int a = 0; a < 3; a++) {
        long long np = (pos[a] / rg) * rg;
        if (pos[a] < np) np -= rg;
        out[a] = int(np);
    }
}

int Volume::determine_depth(int root_pos[3]) {
    // DetermineDepth (gvdb_volume_gvdb.cpp:821-837)
    int pmin[3] = {int(posMin_[0]), int(posMin_[1]), int(posMin_[2])};
    int pmax[3] = {int(posMax_[0]), int(posMax_[1]), int(posMax_[2])};
    for (int lev = 0; lev < levels_; lev++) {
        int a[3], b[3];
        covering_node_host(lev, pmin, a);
        covering_node_host(lev, pmax, b);
        if (a[0] == b[0] && a[1] == b[1] && a[2] == b[2]) {
            root_pos[0] = a[0]; root_pos[1] = a[1]; root_pos[2] = a[2];
            return lev;
        }
    }
    return -1;
}

// --------------------------------------------------------- GPU topology ----
int Volume::rebuild_topology(int n, float radius, const float orig[3]) {
    Volume* v = this;
    return multi_rebuild_topology(&v, 1, &n, &radius, orig, self_exec(), nullptr);
}

// -------------------------------------------------------- host topology ----
bool Volume::pos_in_node(u64 nodeid, const int pos[3], uint32_t& bit) {
    // getPosInNode (gvdb_volume_gvdb.cpp:2858-2875)
    Node* c = hnode(nodeid);
    int lev = c->mLev;
    long long rg = range_[lev];
    int res = res_[lev];
    long long p[3] = {(long long)pos[0] - c->mPos[0], (long long)pos[1] - c->mPos[1], (long long)pos[2] - c->mPos[2]};
    if (p[0] >= 0 && p[1] >= 0 && p[2] >= 0 && p[0] < rg && p[1] < rg && p[2] < rg) {
        for (int a = 0; a < 3; a++) p[a] = (p[a] * res) / rg;
        bit = uint32_t((p[2] * res + p[1]) * res + p[0]);
        return true;
    }
    bit = 0;
    return false;
}

u64 Volume::child_node_host(u64 nodeid, uint32_t b) {
    Node* c = hnode(nodeid);
    if (c->mChildList == ID_UNDEFL) return ID_UNDEF64;
    return hclist(c->mChildList)[b];
}

u64 Volume::insert_child(u64 nodeid, u64 childid, uint32_t i) {
    // InsertChild, dense child list (gvdb_volume_gvdb.cpp:2980-3027)
    hnode(childid)->mParent = nodeid;
    if (hnode(nodeid)->mChildList == ID_UNDEFL) {
        int lev = hnode(nodeid)->mLev;
        u64 cl = pool_alloc(1, lev);
        hnode(nodeid)->mChildList = cl;
        u64* l = hclist(cl);
        for (u64 k = 0; k < voxcnt(lev); k++) l[k] = ID_UNDEF64;
    }
    hclist(hnode(nodeid)->mChildList)[i] = childid;
    return childid;
}

u64 Volume::reparent(int lev, u64 prevroot, const int pos[3], bool& bnew) {
    // Reparent (gvdb_volume_gvdb.cpp:2726-2763)
    int pp[3] = {hnode(prevroot)->mPos[0], hnode(prevroot)->mPos[1], hnode(prevroot)->mPos[2]};
    int p1[3] = {0, 0, 0}, p2[3];
    bool cover = false;
    while (!cover) {
        lev++;
        if (lev >= levels_) return ID_UNDEFL;
        covering_node_host(lev, pos, p1);
        covering_node_host(lev, pp, p2);
        cover = (p1[0] == p2[0] && p1[1] == p2[1] && p1[2] == p2[2]);
    }
    u64 nr = pool_alloc(0, lev);
    if (nr == ID_UNDEFL) return ID_UNDEFL;
    setup_node(nr, lev, p1);
    bool bn = false;
    activate_space_r(nr, pp, bn, prevroot, 0);
    u64 leaf = activate_space_r(nr, pos, bnew, ID_UNDEFL, 0);
    root_ = nr;
    if (leaf == ID_UNDEFL) return ID_UNDEFL;
    return hnode(leaf)->mParent;
}

u64 Volume::activate_space_r(u64 nodeid, const int pos[3], bool& bnew, u64 stopnode, int stoplev) {
    // ActivateSpace (gvdb_volume_gvdb.cpp:2804-2855)
    if (root_ == ID_UNDEFL && nodeid == root_) {
        int p[3];
        covering_node_host(stoplev, pos, p);
        root_ = pool_alloc(0, stoplev);
        setup_node(root_, stoplev, p);
        nodeid = root_;
    }
    uint32_t b;
    if (pos_in_node(nodeid, pos, b)) {
        Node* cur = hnode(nodeid);
        if (stopnode != ID_UNDEFL) {
            Node* sn = hnode(stopnode);
            if (pos[0] == sn->mPos[0] && pos[1] == sn->mPos[1] && pos[2] == sn->mPos[2] &&
                child_node_host(nodeid, b) == ID_UNDEF64 && cur->mLev == sn->mLev + 1)
                return insert_child(nodeid, stopnode, b);
        }
        if (cur->mLev == stoplev) return nodeid;
        u64 child;
        if (child_node_host(nodeid, b) == ID_UNDEF64) {
            int lev = cur->mLev;
            int ppos[3] = {cur->mPos[0], cur->mPos[1], cur->mPos[2]};
            // AddChildNode (:2707-2723) with getPosFromBit
            child = pool_alloc(0, lev - 1);
            long long rg = range_[lev - 1];
            int lg = logdim_[lev];
            uint32_t m = (1u << lg) - 1;
            int lp[3] = {int(b & m), int((b >> lg) & m), int((b >> (2 * lg)) & m)};
            int cp[3] = {int(lp[0] * rg + ppos[0]), int(lp[1] * rg + ppos[1]), int(lp[2] * rg + ppos[2])};
            setup_node(child, lev - 1, cp);
            insert_child(nodeid, child, b);
            if (lev == 1) bnew = true;
        } else {
            child = child_node_host(nodeid, b);
        }
        if (((child >> 8) & 0xFF) == 0) return child;
        return activate_space_r(child, pos, bnew, stopnode, stoplev);
    }
    u64 parent = hnode(nodeid)->mParent;
    if (parent == ID_UNDEFL) {
        parent = reparent(hnode(nodeid)->mLev, nodeid, pos, bnew);
        if (parent == ID_UNDEFL) return ID_UNDEFL;
    }
    return activate_space_r(parent, pos, bnew, stopnode, stoplev);
}

int Volume::rebuild_topology_cpu(int n, const float orig[3], const float* host) {
    if (levels_ == 0) return fail(GVDB_ERR_STATE, "RebuildTopologyCPU before Configure");
    if (n < 0 || (n > 0 && !host)) return fail(GVDB_ERR_INVALID, "RebuildTopologyCPU: null positions");
    int rc = sync_host();
    if (rc) return rc;
    for (int i = 0; i < n; i++) {
        float p[3] = {host[3 * i] + orig[0], host[3 * i + 1] + orig[1], host[3 * i + 2] + orig[2]};
        int ip[3] = {int(p[0]), int(p[1]), int(p[2])};
        bool bnew = false;
        activate_space_r(root_, ip, bnew, ID_UNDEFL, 0);
    }
    dev_stale_ = true;
    topo_gen_++;
    compute_top_lev();
    return GVDB_OK;
}

int Volume::activate_space(const float pos[3]) {
    if (levels_ == 0) return fail(GVDB_ERR_STATE, "ActivateSpace before Configure");
    int rc = sync_host();
    if (rc) return rc;
    int ip[3] = {int(pos[0]), int(pos[1]), int(pos[2])};
    bool bnew = false;
    activate_space_r(root_, ip, bnew, ID_UNDEFL, 0);
    dev_stale_ = true;
    topo_gen_++;
    compute_top_lev();
    return GVDB_OK;
}

// ----------------------------------------------------- finish and atlas ----
int Volume::finish_topology(bool commit, bool bounds) {
    Volume* v = this;
    return multi_finish_topology(&v, 1, commit, bounds, self_exec(), nullptr);
}

int Volume::resize_atlas(int cz) {
    // AtlasResize: Z grows, content kept (gvdb_allocator.cpp:622-650)
    if (cz == atlas_cnt_[2]) return GVDB_OK;
    int br = brickwid_ + 2 * apron_;
    u64 slots = u64(atlas_cnt_[0]) * atlas_cnt_[1] * cz;
    for (auto& ch : chan_) {
        if (!ch.exists) continue;
        size_t old = size_t(atlas_max_) * br * br * br * ch.elem();
        size_t need = size_t(slots) * br * br * br * ch.elem();
        if (need > ch.data.bytes) {
            DevBuf nb;
            HIPCHK(nb.ensure(need));
            HIPCHK(launch_fill_u32(nb.as<uint32_t>(), 0u, nb.bytes / 4, stream));
            if (old) HIPCHK(hipMemcpyAsync(nb.p, ch.data.p, std::min(old, need), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipStreamSynchronize(stream));
            ch.data.release();
            ch.data = nb;
        }
    }
    atlas_cnt_[2] = cz;
    atlas_max_ = slots;
    atlas_epoch_++;
    return GVDB_OK;
}

Tree Volume::make_tree() {
    Tree t;
    memset(&t, 0, sizeof(t));
    t.levels = levels_;
    t.top_lev = top_lev_;
    for (int l = 0; l < levels_; l++) {
        t.logdim[l] = logdim_[l];
        t.res[l] = res_[l];
        t.range[l] = int(range_[l]);
        t.nodecnt[l] = int(pool_[0][l].lastEle);
        t.nodes[l] = pool_[0][l].dev.as<Node>();
        t.clists[l] = l > 0 ? pool_[1][l].dev.as<u64>() : nullptr;
    }
    t.amap = amap_.as<AtlasNode>();
    for (int a = 0; a < 3; a++) t.atlas_cnt[a] = atlas_cnt_[a];
    t.apron = apron_;
    t.brickwid = brickwid_;
    t.brick_res = brickwid_ + 2 * apron_;
    t.brick_vox = t.brick_res * t.brick_res * t.brick_res;
    for (int a = 0; a < 3; a++) { t.bmin[a] = objMin_[a]; t.bmax[a] = objMax_[a]; }
    t.epsilon = epsilon_;
    return t;
}

int Volume::update_atlas() {
    Volume* v = this;
    return multi_update_atlas(&v, 1, self_exec(), nullptr);
}

int Volume::clear_channel(int c) {
    Volume* v = this;
    return multi_clear_channel(&v, 1, c, self_exec(), nullptr);
}

// ------------------------------------------------------ points -> voxels ----
int Volume::insert_points_subcell(int scs, int n, float radius, const float trans[3], int* len) {
    Volume* v = this;
    return multi_insert(&v, 1, scs, &n, &radius, trans, len, self_exec(), nullptr);
}

int Volume::gather_density(int scs, int n, float radius, const float trans[3], int len, int chanD, int chanC, bool accum) {
    (void)n; (void)trans;
    Volume* v = this;
    return multi_gather(&v, 1, scs, &len, &radius, chanD, accum, false, self_exec(), nullptr, false, chanC);
}

// GatherLevelSet (gvdb_volume_gvdb.cpp:5348-5371)
int Volume::gather_level_set(int scs, int n, float radius, const float trans[3], int len, int chanD, int chanC, bool accum) {
    (void)n; (void)trans;
    Volume* v = this;
    return multi_gather(&v, 1, scs, &len, &radius, chanD, accum, true, self_exec(), nullptr, false, chanC);
}

// InsertPointsSubcell_FP16 (gvdb_volume_gvdb.cpp:5094-5189): the subcell
// tables of InsertPointsSubcell with the FP16 binning rule, the entries
// quantised to u16 against the point bounds of the last RebuildTopology
// (mPosMin, mPosRange: :5452-5462).
int Volume::insert_points_subcell_fp16(int scs, int n, float radius, const float trans[3], int* len) {
    float mn[3], rg[3];
    for (int a = 0; a < 3; a++) {
        mn[a] = posMin_[a];
        rg[a] = posMax_[a] - posMin_[a];
        if (!(rg[a] > 0.0f))
            return fail(GVDB_ERR_STATE, "InsertPointsSubcell_FP16: no point bounds (call RebuildTopology first)");
    }
    Volume* v = this;
    int rc = multi_insert(&v, 1, scs, &n, &radius, trans, len, self_exec(), nullptr, BIN_FP16);
    if (rc) return rc;
    if (entries_stale_) {
        rc = multi_materialize(&v, 1, self_exec(), nullptr);
        if (rc) return rc;
    }
    const u64 n3 = u64(sc_len_) * 3;
    HIPCHK(sc_pnt16_.ensure(size_t(n3) * 2));
    HIPCHK(launch_quantize_entries(sc_pnt_.as<float>(), sc_pnt16_.as<uint16_t>(), n3, mn, rg, stream));
    for (int a = 0; a < 3; a++) { q_min_[a] = mn[a]; q_range_[a] = rg[a]; }
    sc16_valid_ = true;
    return GVDB_OK;
}

// GatherLevelSet_FP16 (gvdb_volume_gvdb.cpp:5374-5401): no accumulation.
int Volume::gather_level_set_fp16(int scs, int n, float radius, const float trans[3], int len, int chanD, int chanC) {
    (void)n; (void)trans; (void)chanC;
    if (!sc16_valid_) return fail(GVDB_ERR_STATE, "GatherLevelSet_FP16: call InsertPointsSubcell_FP16 first");
    Volume* v = this;
    return multi_gather(&v, 1, scs, &len, &radius, chanD, false, true, self_exec(), nullptr, true);
}

// Compute (gvdb_volume_gvdb.cpp:4527-4552): `iterations` launches of the
// effect, each followed by UpdateApron(chan, boundval) when asked.
int Volume::compute(int effect, int c, int iterations, const float p[3], bool update_apron_each, bool skip_aprons,
                    float boundval) {
    if (c < 0 || c >= 8 || !chan_[c].exists) return fail(GVDB_ERR_INVALID, "Compute: no such channel");
    if (chan_[c].type == T_UCHAR4)
        return fail(GVDB_ERR_UNSUPPORTED, "Compute: the uchar effects are not provided (float and half channels)");
    if (effect != OP_FILL && effect != OP_SMOOTH && effect != OP_NOISE && effect != OP_GROW)
        return fail(GVDB_ERR_UNSUPPORTED, "Compute: effect not provided (float channels: fill, smooth, noise, grow)");
    if (effect != OP_FILL && apron_ != 1)
        return fail(GVDB_ERR_UNSUPPORTED, "Compute: smooth, noise and grow need an apron of 1");
    if (iterations < 0) return fail(GVDB_ERR_INVALID, "Compute: negative iteration count");
    int rc = commit_pools();
    if (rc) return rc;
    ComputeArgs a;
    memset(&a, 0, sizeof(a));
    a.t = make_tree();
    a.atlas = chan_[c].data.p;
    a.atlas_type = chan_[c].type;
    a.op = effect;
    a.p1 = p[0]; a.p2 = p[1]; a.p3 = p[2];
    a.slots = int(std::min<u64>(atlas_used_, atlas_max_));
    const int br = brickwid_ + 2 * apron_;
    for (int k = 0; k < 3; k++) {
        // the reference's unpacked index b * 8 + t + 1 over ceil(count / 8) blocks
        const int res = atlas_cnt_[k] * br;
        const int count = skip_aprons ? atlas_cnt_[k] * brickwid_ : res;
        const int blocks = (count + 7) / 8;
        a.lo[k] = 1;
        a.hi[k] = std::min(8 * blocks, res - 1);
    }
    for (int it = 0; it < iterations; it++) {
        HIPCHK(launch_compute(a, stream));
        atlas_epoch_++;
        if (update_apron_each) {
            rc = update_apron(c, boundval);
            if (rc) return rc;
        }
    }
    return GVDB_OK;
}

int Volume::update_apron(int c, float boundval) {
    Volume* v = this;
    return multi_apron(&v, 1, c, boundval, self_exec(), nullptr);
}

// ---------------------------------------------------------------- render ----
int Volume::linear_tf(float t0, float t1, const float a[4], const float b[4]) {
    // Scene::LinearTransferFunc (gvdb_scene.cpp:66-88)
    int n0 = int(t0 * float(TF_SIZE)), n1 = int(t1 * float(TF_SIZE));
    n0 = std::max(0, n0);
    n1 = std::min(TF_SIZE, n1);
    for (int k = n0; k < n1; k++) {
        float x = float(k - n0) / float(n1 - n0);
        for (int c = 0; c < 4; c++) tf_host[4 * k + c] = a[c] + x * (b[c] - a[c]);
    }
    tf_committed = false;
    return GVDB_OK;
}

int Volume::commit_tf() {
    if (capture_) return not_recordable("committing the transfer function (render once before recording)");
    HIPCHK(tf_dev.ensure(size_t(TF_SIZE) * 16));
    HIPCHK(hipMemcpyAsync(tf_dev.p, tf_host.data(), size_t(TF_SIZE) * 16, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    tf_committed = true;
    return GVDB_OK;
}

int Volume::add_render_buf(int buf, int w, int h, int bpp) {
    if (buf < 0 || buf >= 4 || w <= 0 || h <= 0 || (bpp != 4 && bpp != 16))
        return fail(GVDB_ERR_INVALID, "AddRenderBuf: buffer 0..3, positive size, 4 or 16 bytes per pixel");
    RenderBuf& r = rbuf_[buf];
    r.w = w; r.h = h; r.bpp = bpp;
    HIPCHK(r.data.ensure(size_t(w) * h * bpp));
    return GVDB_OK;
}

int Volume::render(int shading, int c, int buf) {
    Volume* v = this;
    return multi_render(&v, 1, &shading, &c, &buf, self_exec(), nullptr);
}

int Volume::read_render_buf(int buf, void* dst) {
    if (buf < 0 || buf >= 4 || rbuf_[buf].w == 0 || !dst) return fail(GVDB_ERR_INVALID, "ReadRenderBuf: no such buffer");
    RenderBuf& r = rbuf_[buf];
    HIPCHK(hipMemcpyAsync(dst, r.data.p, size_t(r.w) * r.h * r.bpp, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return GVDB_OK;
}

unsigned long long Volume::device_bytes() const {
    unsigned long long b = 0;
    for (int g = 0; g < 2; g++)
        for (int l = 0; l < MAXLEV; l++) b += pool_[g][l].dev.bytes;
    for (const auto& c : chan_) b += c.data.bytes;
    for (const auto& r : rbuf_) b += r.data.bytes;
    const DevBuf* bufs[] = {&amap_, &nbr_, &nmask_, &nbase_, &faces_, &lenparts_, &pts_owned_, &sc_cnt_, &sc_off_, &sc_pos_, &sc_nid_, &sc_map_, &sc_pnt_,
                            &sc_use_, &bkt_key_, &bkt_slot_, &bkt_cnt_, &bkt_off_, &sorted_, &sc_flags_, &small_,
                            &masks_, &bitmaps_, &ranks_, &scan_tmp_, &flags_, &slot_scan_, &tf_dev,
                      &incr_marker_, &sc_pnt16_, &view_dev_, &view_scene_, &kb_keys_, &kb_keysB_, &kb_cnt_, &kb_off_,
                      &kb_tmp_, &kb_hist_, &kb_flags_, &kb_pos_, &kb_list_, &btab_, &bbox_part_};
    for (const DevBuf* d : bufs) b += d->bytes;
    return b;
}

// ---------------------------------------------------------------- access ----
int Volume::get_pool(int grp, int lev, void* dst, u64* count, u64* stride) {
    if (grp < 0 || grp > 1 || lev < 0 || lev >= levels_) return fail(GVDB_ERR_INVALID, "get_pool: bad group or level");
    Pool& p = pool_[grp][lev];
    if (count) *count = p.exists ? p.lastEle : 0;
    if (stride) *stride = p.stride;
    if (!dst || !p.exists || !p.lastEle) return GVDB_OK;
    int rc = sync_host();
    if (rc) return rc;
    memcpy(dst, p.host.data(), p.lastEle * p.stride);
    return GVDB_OK;
}

int Volume::resolve_bounds() {
    if (!bounds_pending_) return GVDB_OK;
    HIPCHK(hipEventSynchronize(bounds_ev_));
    const int* r = reinterpret_cast<const int*>(pinned_ + PIN_BOUNDS);
    for (int a = 0; a < 3; a++) { objMin_[a] = float(r[a]); objMax_[a] = float(r[3 + a]); }
    bounds_pending_ = false;
    return GVDB_OK;
}

int Volume::get_info(int* out2, float* outf) {
    int rc = resolve_bounds();
    if (rc) return rc;
    if (out2) { out2[0] = top_lev_; out2[1] = curr_depth_; }
    if (outf)
        for (int a = 0; a < 3; a++) {
            outf[a] = objMin_[a]; outf[3 + a] = objMax_[a]; outf[6 + a] = posMin_[a]; outf[9 + a] = posMax_[a];
        }
    return GVDB_OK;
}

int Volume::get_atlas_info(int c, int* out) {
    if (c < 0 || c >= 8 || !chan_[c].exists) return fail(GVDB_ERR_INVALID, "atlas info: no such channel");
    int br = brickwid_ + 2 * apron_;
    for (int a = 0; a < 3; a++) { out[a] = atlas_cnt_[a]; out[3 + a] = atlas_cnt_[a] * br; }
    out[6] = int(atlas_used_);
    return GVDB_OK;
}

int Volume::read_atlas(int c, void* dst) {
    // Device brick-major storage -> GVDB's 3D atlas layout (x fastest)
    if (c < 0 || c >= 8 || !chan_[c].exists || !dst) return fail(GVDB_ERR_INVALID, "read_atlas: no such channel");
    Channel& ch = chan_[c];
    int br = brickwid_ + 2 * apron_;
    u64 slots = atlas_max_;
    size_t bv = size_t(br) * br * br;
    std::vector<uint8_t> tmp(size_t(slots) * bv * ch.elem());
    HIPCHK(hipMemcpyAsync(tmp.data(), ch.data.p, tmp.size(), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    int rx = atlas_cnt_[0] * br, ry = atlas_cnt_[1] * br;
    int e = ch.elem();
    uint8_t* out = reinterpret_cast<uint8_t*>(dst);
    for (u64 s = 0; s < slots; s++) {
        int sx = int(s % atlas_cnt_[0]), sy = int((s / atlas_cnt_[0]) % atlas_cnt_[1]), sz = int(s / (u64(atlas_cnt_[0]) * atlas_cnt_[1]));
        for (int z = 0; z < br; z++)
            for (int y = 0; y < br; y++) {
                size_t src = (s * bv + size_t(z * br + y) * br) * e;
                size_t d = ((size_t(sz * br + z) * ry + (sy * br + y)) * rx + sx * br) * e;
                memcpy(out + d, tmp.data() + src, size_t(br) * e);
            }
    }
    return GVDB_OK;
}

int Volume::read_atlas_map(void* dst) {
    if (!dst) return fail(GVDB_ERR_INVALID, "read_atlas_map: null destination");
    if (!amap_max_) return GVDB_OK;
    HIPCHK(hipMemcpyAsync(dst, amap_.p, size_t(amap_max_) * sizeof(AtlasNode), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return GVDB_OK;
}

int Volume::read_subcells(int which, void* dst, u64* count) {
    if (entries_stale_ && dst) {
        Volume* v = this;
        int rc = multi_materialize(&v, 1, self_exec(), nullptr);
        if (rc) return rc;
    }
    size_t n = 0, esz = 4;
    DevBuf* b = nullptr;
    switch (which) {
        case 0: b = &sc_cnt_; n = num_sc_; break;
        case 1: b = &sc_off_; n = num_sc_; break;
        case 2: b = &sc_pos_; n = num_sc_; esz = 12; break;
        case 3: b = &sc_nid_; n = num_sc_; break;
        case 4: b = &sc_pnt_; n = sc_len_; esz = 12; break;
        case 5: b = &sc_map_; n = sc_map_n_; break;
        case 6:
            if (!sc16_valid_) return fail(GVDB_ERR_STATE, "read_subcells: no InsertPointsSubcell_FP16 entries");
            b = &sc_pnt16_; n = sc_len_; esz = 6; break;
        default: return fail(GVDB_ERR_INVALID, "read_subcells: which must be 0..6");
    }
    if (count) *count = n;
    if (!dst || !n) return GVDB_OK;
    HIPCHK(hipMemcpyAsync(dst, b->p, n * esz, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return GVDB_OK;
}

}  // namespace gvdb
