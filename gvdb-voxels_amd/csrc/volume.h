// volume.h -- host runtime of one grid: pools, atlas, aux buffers, scene.
#pragma once
#include <hip/hip_runtime.h>
#include <memory>
#include <string>
#include <vector>
#include "gvdb_device.h"
#include "kernels.h"

namespace gvdb {

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n, bool preserve = false, hipStream_t s = nullptr);
    void release();
    template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

struct Pool {
    bool exists = false;
    u64 stride = 0;      // bytes per element
    u64 max = 0;         // capacity in elements (doubling growth, as the reference's PoolAlloc)
    u64 lastEle = 0;
    u64 usedNum = 0;
    std::vector<uint8_t> host;   // host mirror (host-built topology, readback)
    DevBuf dev;
};

struct Channel {
    bool exists = false;
    int type = T_FLOAT;
    int filter = 0, border = 0;
    DevBuf data;
    int elem() const { return type == T_HALF ? 2 : 4; }
};

struct RenderBuf {
    int w = 0, h = 0, bpp = 4;
    DevBuf data;
};

struct CameraParams {
    float angs[3] = {0, 45, 0};
    float to[3] = {0, 0, 0};
    float dist = 120.f, dolly = 1.f, fov = 40.f, znear = 0.1f, zfar = 5000.f;
};

enum Stage { ST_TOPO = 0, ST_ATLAS, ST_INSERT, ST_GATHER, ST_APRON, ST_RENDER, ST_CLEAR, ST_COUNT };

// A stream shared by the grids that run on it; destroyed (after draining)
// when its last user lets go, so a batch may go away before its grids.
struct SharedStream {
    hipStream_t s = nullptr;
    int device = 0;
    ~SharedStream();
    static std::shared_ptr<SharedStream> make(int device, hipError_t& e);
};

// Where the stages of one call run: the stream, argument staging, and a
// SMALL_WORDS-word slot per grid (device and pinned host) for the few values
// a stage reads back (bounding box, node counts, subcell totals).
static const int SMALL_WORDS = 64;
// small per-grid result words (device dsmall / pinned hsmall, SMALL_WORDS each)
enum { SM_BBOX = 0, SM_LINKED = 8, SM_TOTAL = 9, SM_ORPHANS = 10, SM_BUCKETS = 11, SM_LEN = 12, SM_COUNTS = 16, SM_BOUNDS = 40,
       SM_UNFLAGGED = 48 };
struct Graph;
struct Exec {
    hipStream_t s;
    Arena* ar;
    uint32_t* dsmall;
    uint32_t* hsmall;
    Graph* cap = nullptr;   // recording into this graph (graph.h): no host round trips
};

class Volume {
public:
    explicit Volume(int device);
    ~Volume();
    int init();

    std::string err;
    int fail(int code, const std::string& msg) { err = msg; return code; }

    // configuration
    int configure(int levels, const int* log2dims, const int* initcnt);
    int set_channel_default(int x, int y, int z);
    int add_channel(int chan, int dtype, int apron, int filter, int border, int ax, int ay, int az);
    int set_epsilon(float eps, int maxiter);

    // points
    int set_points_gpu(const void* p, int n, int off, int stride);
    int upload_points(const float* host, int n);
    // point colours (SetPoints' colour buffer) and the render's colour channel
    int set_colors_gpu(const void* p, int n, int off, int stride);
    int upload_colors(const uint32_t* host, int n);
    int set_color_channel(int chan);
    int color_channel() const { return clr_chan_; }

    // topology
    // RequestFullRebuild (gvdb_volume_gvdb.h:432): sets mRebuildTopo to `on`
    int request_full_rebuild(bool on) { rebuild_topo_ = on; return 0; }
    int rebuild_topology(int n, float radius, const float orig[3]);
    int accumulate_topology(int n, float radius, const float orig[3], int min_depth);
    int compute(int effect, int chan, int iterations, const float params[3], bool update_apron, bool skip_aprons,
                float boundval);
    // VBX files (vbx.cpp)
    int save_vbx(const char* path);
    int load_vbx(const char* path, int force_major, int force_minor);
    int write_atlas(int chan, const void* src);
    // user kernels (plugin.cpp)
    int compute_kernel(void* hip_function, int chan, bool update_apron, bool skip_aprons);
    int render_kernel(void* hip_function, int chan, int rbuf);
    int upload_view(int chan, int rbuf, void** dev);
    void get_transform(float out12[12]) const;
    bool identity_transform() const;
    // SetTransform (gvdb_volume_gvdb.cpp:5770-5793): the grid's placement in
    // the application's space, xform = T R S PT; the march takes rays and the
    // light back into index space through its inverse
    int set_transform(const float pretrans[3], const float scale[3], const float angs[3], const float trans[3]);
    int rebuild_topology_cpu(int n, const float orig[3], const float* host);
    int activate_space(const float pos[3]);
    int clear();
    int finish_topology(bool commit, bool bounds);
    int update_atlas();
    int clear_channel(int chan);

    // points -> voxels
    int insert_points_subcell(int scs, int n, float radius, const float trans[3], int* len);
    int gather_density(int scs, int n, float radius, const float trans[3], int len, int chanD, int chanC, bool accum);
    int gather_level_set(int scs, int n, float radius, const float trans[3], int len, int chanL, int chanC, bool accum);
    int insert_points_subcell_fp16(int scs, int n, float radius, const float trans[3], int* len);
    int gather_level_set_fp16(int scs, int n, float radius, const float trans[3], int len, int chanL, int chanC);
    int update_apron(int chan, float boundval);
    // UpdateApron with the neighbour table in the library's fixed device
    // array (no allocation, at most NOALLOC_MAX_BRICKS bricks)
    int update_apron_noalloc(int chan, float boundval);

    // render
    CameraParams cam;
    float light_[3] = {0, 84.85f, 84.85f};
    float back_[4] = {0, 0, 0, 1};
    float extinct_[3] = {-1.1f, 1.5f, 0.f}, steps_[3] = {1.f, 16.f, 0.1f}, cutoff_[3] = {0.005f, 0.01f, 0.f};
    float vrange_[3] = {0.1f, 0.f, 1.f}, shadow_[3] = {0.8f, 1.f, 0.f};
    std::vector<float> tf_host;
    DevBuf tf_dev;
    bool tf_committed = false;
    int linear_tf(float t0, float t1, const float a[4], const float b[4]);
    int commit_tf();
    int add_render_buf(int buf, int w, int h, int bpp);
    int render(int shading, int chan, int buf);
    int read_render_buf(int buf, void* dst);
    void* render_buf_ptr(int buf) const { return (buf >= 0 && buf < 4) ? rbuf_[buf].data.p : nullptr; }
    unsigned long long device_bytes() const;
    void scene_block(int w, int h, ScnBlock& s) const;

    // access
    int get_pool(int grp, int lev, void* dst, u64* count, u64* stride);
    int get_info(int* out2, float* outf);
    int get_atlas_info(int chan, int* out7);
    int read_atlas(int chan, void* dst);
    int read_atlas_map(void* dst);
    int read_subcells(int which, void* dst, u64* count);
    int stage_ms(float* out);
    int kernel_ms(float* out);

    hipStream_t stream = nullptr;
    int device = 0;
    Exec self_exec();

    // HIP graph capture (graph.cpp): while capture() is set, the calls of the
    // per-frame chain are recorded into it instead of run
    bool capturing() const { return capture_ != nullptr; }
    Graph* capture() const { return capture_; }
    void set_capture(Graph* g) { capture_ = g; }
    // fails the recording with `code` (a call that cannot be recorded)
    int not_recordable(const char* what);
    std::shared_ptr<SharedStream> stream_owner() const { return cur_stream_; }
    unsigned topo_generation() const { return topo_gen_; }
    void remember_graph(Graph* g) { graphs_.push_back(g); }
    void forget_graph(Graph* g);
    // after a replay: the recorded insert's readback (entry count, points
    // it could not place) into this grid's state
    int finish_recorded_insert(const uint32_t* slot);
    // moves this grid's work to another stream (null: its own), ordered
    // after what is already queued, without waiting on the host
    int use_stream(const std::shared_ptr<SharedStream>& s);

    // Stages over G grids at once (pipeline.cpp); the single-grid methods
    // above call these with G = 1.  On failure `failed` names the grid.
    static int multi_rebuild_topology(Volume* const* v, int G, const int* n, const float* radius, const float* orig3,
                                      const Exec& ex, int* failed);
    static int multi_full_build(Volume* const* v, int G, const int* n, const float* radius, const float* orig3,
                                const int* depth, const int* rootPos, const float* mnmx, const Exec& ex, int* failed);
    int incremental_topology(int n, float radius, const float orig[3], int D, bool accum, const float* mnmx,
                             const Exec& ex);
    // the per-level brick boxes of a full build (false: level-0 coordinates
    // beyond 16 bits), and the full build over sorted brick keys for boxes
    // too large for bitmaps (keybuild.cpp)
    static bool plan_domains(const Volume* v, int D, const int rootPos[3], const float* mn, const float* mx,
                             float radius, BitDomain* dom);
    int keyed_build(int n, float radius, const float orig[3], int D, const int rootPos[3], const BitDomain* dom,
                    const Exec& ex);
    int kb_sort_unique(u64* keys, u64 cnt, int bits, u64 sentinel, DevBuf& out, uint32_t* count, const Exec& ex);
    void kb_release();
    static int multi_finish_topology(Volume* const* v, int G, bool commit, bool bounds, const Exec& ex, int* failed);
    static int multi_update_atlas(Volume* const* v, int G, const Exec& ex, int* failed);
    static int multi_clear_channel(Volume* const* v, int G, int chan, const Exec& ex, int* failed);
    static int multi_insert(Volume* const* v, int G, int scs, const int* n, const float* radius, const float* trans3,
                            int* len, const Exec& ex, int* failed, int rule = 0);
    // levelset: GatherLevelSet; fp16 (with levelset): GatherLevelSet_FP16 over
    // the u16 entries of InsertPointsSubcell_FP16
    // chanC >= 0: also the colour channel, from the colours the insert carried
    static int multi_gather(Volume* const* v, int G, int scs, const int* len, const float* radius, int chanD, bool accum,
                            bool levelset, const Exec& ex, int* failed, bool fp16 = false, int chanC = -1);
    static int multi_materialize(Volume* const* v, int G, const Exec& ex, int* failed);
    static int multi_apron(Volume* const* v, int G, int chan, float boundval, const Exec& ex, int* failed);
    static int multi_render(Volume* const* v, int G, const int* shading, const int* chan, const int* buf, const Exec& ex,
                            int* failed);
    int hip_fail(hipError_t e, const char* what);

private:
    // pools
    int levels_ = 0;
    int logdim_[MAXLEV] = {0}, res_[MAXLEV] = {0};
    long long range_[MAXLEV] = {0};
    Pool pool_[2][MAXLEV];
    u64 root_ = ID_UNDEFL;
    bool rebuild_topo_ = true;
    int curr_depth_ = -1;
    int top_lev_ = 1;
    bool host_valid_ = true;   // host mirrors hold the current pools
    bool dev_stale_ = false;   // device pools are older than the host mirrors
    unsigned topo_gen_ = 0;    // bumps on every topology change
    unsigned atlas_gen_ = ~0u; // topology generation UpdateAtlas last ran on
    unsigned sc_gen_ = ~0u;    // topology generation InsertPointsSubcell last ran on
    float posMin_[3] = {0, 0, 0}, posMax_[3] = {0, 0, 0};
    // grid transform kept for VBX files (SetTransform's defaults)
    float xf_pretrans_[3] = {0, 0, 0}, xf_angs_[3] = {0, 0, 0}, xf_scale_[3] = {1, 1, 1}, xf_trans_[3] = {0, 0, 0};
    // column-major 4x4: inverse of xform, and inverse of its linear part (R S)
    float xf_inv_[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float xf_invrot_[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float objMin_[3] = {0, 0, 0}, objMax_[3] = {0, 0, 0};
    // FinishTopology's device bounds arrive in pinned_[PIN_BOUNDS..+6) after
    // bounds_ev_; objMin_/objMax_ take them when something reads them
    // (resolve_bounds: the render, the user-kernel view, GetInfo)
    static const int PIN_BOUNDS = 192;
    bool bounds_pending_ = false;
    hipEvent_t bounds_ev_ = nullptr;
    hipEvent_t small_ev_ = nullptr;   // InsertPointsSubcell's early readback of its small words
    int resolve_bounds();
    float epsilon_ = 0.001f;
    int maxiter_ = 256;

    // atlas
    int default_axiscnt_[3] = {16, 16, 1};
    int atlas_cnt_[3] = {0, 0, 0};
    int apron_ = 0;
    int brickwid_ = 0;
    u64 atlas_max_ = 0, atlas_used_ = 0;
    Channel chan_[8];
    DevBuf amap_;
    u64 amap_max_ = 0;
    DevBuf nbr_, nmask_, nbase_, lenparts_;
    // The scatter gather's copy of each brick's boundary faces, current while
    // nothing else has written the atlas (atlas_epoch_ counts atlas writes).
    DevBuf faces_;
    unsigned atlas_epoch_ = 0, faces_epoch_ = ~0u, faces_topo_ = ~0u;
    int faces_chan_ = -1;
    unsigned nbr_gen_ = ~0u;
    void* noalloc_ = nullptr;   // this device's copy of the no-allocation scratch array
    Graph* capture_ = nullptr;      // the graph being recorded, if any
    std::vector<Graph*> graphs_;    // graphs that replay calls on this grid

    // points
    PointSrc pts_{nullptr, 0, 0, 12};
    DevBuf pts_owned_;
    PointSrc clr_{nullptr, 0, 0, 4};   // colours, or base null
    DevBuf clr_owned_;
    int clr_chan_ = CHAN_UNDEF;
    bool bkt_clr_ = false;             // the current buckets carry colours

    // subcells
    DevBuf sc_cnt_, sc_use_, sc_off_, sc_pos_, sc_nid_, sc_map_, sc_pnt_, sc_flags_;
    DevBuf bkt_key_, bkt_slot_, bkt_cnt_, bkt_off_, sorted_;
    int sc_len_ = 0, num_sc_ = 0, sc_bricks_ = 0, sc_map_n_ = 0;
    bool last_insert_pull_ = false;
    bool entries_stale_ = false;   // pull insert done, subcell tables not built yet
    // InsertPointsSubcell_FP16: u16 entries (3 per entry) and the bounds they
    // are quantised against; valid until the next insert
    DevBuf sc_pnt16_;
    bool sc16_valid_ = false;
    float q_min_[3] = {0, 0, 0}, q_range_[3] = {0, 0, 0};
    int bkt_scs_ = 0, bkt_R_ = 0;
    float bkt_radius_ = 0.f;
    unsigned bkt_gen_ = ~0u;       // topology generation of the current buckets

    // topology scratch
    DevBuf small_;       // bbox keys, counts
    DevBuf bbox_part_;   // the bounding box's per-block keys
    DevBuf bitmaps_, ranks_, masks_, scan_tmp_, flags_, slot_scan_;
    DevBuf kb_keys_, kb_keysB_, kb_cnt_, kb_off_, kb_tmp_, kb_hist_, kb_flags_, kb_pos_, kb_list_;   // keyed builds
    DevBuf incr_marker_;   // per-leaf markers of the incremental build
    // the brick table of the last full rebuild (BrickTab), valid while
    // btab_gen_ equals topo_gen_
    DevBuf btab_;
    BitDomain btab_dom_{};
    unsigned btab_gen_ = ~0u;
    BrickTab brick_tab() const {
        if (btab_gen_ != topo_gen_ || !btab_.p) return BrickTab{nullptr, {0, 0, 0}, {0, 0, 0}};
        const BitDomain& d = btab_dom_;
        return BrickTab{btab_.as<u64>(), {d.lo[0], d.lo[1], d.lo[2]}, {d.dim[0], d.dim[1], d.dim[2]}};
    }
    DevBuf view_dev_, view_scene_;   // the device view handed to user kernels
    uint32_t* pinned_ = nullptr;   // host staging (256 words)
    Arena arena_;
    std::shared_ptr<SharedStream> own_stream_, cur_stream_;
    hipEvent_t handoff_ = nullptr;

    // render
    RenderBuf rbuf_[4];

    // timing
    hipEvent_t ev_[ST_COUNT][2] = {};
    bool ev_used_[ST_COUNT] = {};
    hipEvent_t kev_[4][2] = {};
    bool kev_used_[4] = {};

    void stage_begin(Stage s);
    void stage_end(Stage s);
    u64 pool_alloc(int grp, int lev);
    int ensure_pool_capacity(int grp, int lev, u64 need, bool preserve = false, hipStream_t s = nullptr);
    int sync_host();
    int commit_pools();
    Tree make_tree();
    void setup_node(u64 id, int lev, const int pos[3]);
    void compute_top_lev();
    int determine_depth(int root_pos[3]);
    void covering_node_host(int lev, const int pos[3], int out[3]) const;
    u64 activate_space_r(u64 nodeid, const int pos[3], bool& bnew, u64 stopnode, int stoplev);
    u64 reparent(int lev, u64 prevroot, const int pos[3], bool& bnew);
    u64 insert_child(u64 nodeid, u64 childid, uint32_t i);
    bool pos_in_node(u64 nodeid, const int pos[3], uint32_t& bit);
    u64 child_node_host(u64 nodeid, uint32_t b);
    Node* hnode(u64 ref);
    u64* hclist(u64 ref);
    int resize_atlas(int cz);
    u64 voxcnt(int lev) const { u64 r = u64(1) << logdim_[lev]; return r * r * r; }
};

}  // namespace gvdb
