// kernels.h -- argument records and launchers of the gfx950 kernels.
//
// Every launcher takes G argument records (one per grid) and issues one
// launch per kernel over all of them (multi.h); G = 1 is the single-grid
// form with the record passed by value.  Launchers enqueue on the given
// stream and return the launch status.
#pragma once
#include <hip/hip_runtime.h>
#include "gvdb_device.h"
#include "multi.h"

namespace gvdb {

// Points are read as float3 at (base + i*stride + off), as GVDB's DataPtr
// (src/gvdb_allocator.h:35-55 of the reference) describes them.
struct PointSrc {
    const char* base;
    int n;
    int off;
    int stride;
};

// A box of node coordinates at one level, stored as an x-major bitmap so that
// bit order equals GVDB's sorted key order (lev, x, y, z).
struct BitDomain {
    int lo[3];
    int dim[3];
    u64 nbits;
    u64 nwords;   // 32-bit words
};

// ---- utilities (util.hip)
struct FillArgs {
    void* p;
    u64 n;        // elements
    u64 v;
    int wide;     // 0: u32 elements, 1: u64 elements
};
hipError_t launch_fill_multi(const FillArgs* a, int G, Arena* ar, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, u64 n, hipStream_t s);
hipError_t launch_fill_u64(u64* p, u64 v, u64 n, hipStream_t s);
// *out = the sum of in[0, n) in 64 bits (the exact total behind a u32 scan,
// whose own total wraps past 2^32)
hipError_t launch_sum_u32(const uint32_t* in, u64 n, u64* out, hipStream_t s);

// Exclusive scan of u32 values; `mode` selects what is scanned:
//   0: in0[i];  1: popcount(in0[i]);  2: popcount(in1[i] & ~in0[i]);  3: in0[i] != 0
// `total` (device, may be null) receives the sum; `sums` must hold
// scan_tmp_words(n) words.
struct ScanArgs {
    const uint32_t* in0;
    const uint32_t* in1;
    int mode;
    uint32_t* out;
    u64 n;
    uint32_t* sums;
    uint32_t* total;
};
u64 scan_tmp_words(u64 n);
hipError_t launch_scan_multi(const ScanArgs* a, int G, Arena* ar, hipStream_t s);
hipError_t launch_scan_u32(const uint32_t* in0, const uint32_t* in1, int mode, uint32_t* out, u64 n,
                           uint32_t* total, uint32_t* tmp, hipStream_t s);

// ---- topology (topology.hip)
struct BboxArgs {
    PointSrc pts;
    float o[3];
    uint32_t* keys6;      // orderable float keys: min x,y,z then max x,y,z
    uint32_t* part;       // BBOX_PART_WORDS: six keys per block
    uint32_t nparts;      // set by the launcher: blocks of the grid
};
static const int BBOX_BLOCKS = 2048;                         // blocks per grid at most
static const int BBOX_PART_WORDS = BBOX_BLOCKS * 6;
hipError_t launch_bbox_multi(const BboxArgs* a, int G, Arena* ar, hipStream_t s);

struct MarkArgs {
    PointSrc pts;
    float o[3];
    float radius;
    int dilate;
    int range0;
    float inv_range0;
    BitDomain d0;
    uint32_t* bmA;
    uint32_t* bmB;
    uint32_t* masks;      // optional zeroed u32 per level-0 domain brick
    uint8_t* bytes;       // optional zeroed byte per domain brick: A at 0, B at nwords * 32 (instead of masks)
};
hipError_t launch_mark_points_multi(const MarkArgs* a, int G, Arena* ar, hipStream_t s);

struct ParentArgs {
    BitDomain dc;
    const uint32_t* Ac;
    const uint32_t* Bc;
    int shift;
    BitDomain dp;
    uint32_t* Ap;
    uint32_t* Bp;
};
hipError_t launch_mark_parents_multi(const ParentArgs* a, int G, Arena* ar, hipStream_t s);

struct LevelBuild {
    BitDomain dom;
    const uint32_t* bmA;     // S1 bitmap
    const uint32_t* bmB;     // dilated bitmap (may equal bmA when no dilation)
    const uint32_t* rankA;   // exclusive scan of popcount(bmA)
    const uint32_t* rankE;   // exclusive scan of popcount(bmB & ~bmA)
    uint32_t countA;
    uint32_t countE;
    int lev;
    int range;
};
struct EmitArgs {
    LevelBuild L;
    Node* nodes;
    int has_clist;
    u64* btab;            // optional (level 0, box up to BRICK_TAB_MAX bits): the brick table
};
hipError_t launch_emit_nodes_multi(const EmitArgs* a, int G, Arena* ar, hipStream_t s);

struct LinkArgs {
    LevelBuild C;
    LevelBuild P;            // ignored when parent_is_root
    int parent_is_root;
    int logdim_p;
    Node* cnodes;
    u64* pclists;
    u64 clist_len;
};
hipError_t launch_link_level_multi(const LinkArgs* a, int G, Arena* ar, hipStream_t s);

// ---- keyed topology build (keysort.hip)
struct KeyGenArgs {
    PointSrc pts;
    float o[3];
    float radius;
    int dilate;
    int range0;
    float inv_range0;
    BitDomain d0;
    u64 sentinel;           // own key of a point outside the box (sorts last)
    u64* keysA;             // per point: own brick's key or the sentinel
    uint32_t* cntB;         // per point: bricks of its dilation box in the box
    const uint32_t* offB;   // exclusive scan of cntB (emit pass)
    u64* keysB;
};
hipError_t launch_key_gen(const KeyGenArgs& a, bool emit, hipStream_t s);
u64 radix_hist_words(u64 n);
// sorts n keys of `bits` significant bits (LSD, 8 bits a pass); *result is
// keys or tmp, whichever holds the sorted keys
hipError_t radix_sort_u64(u64* keys, u64* tmp, u64 n, int bits, uint32_t* hist, uint32_t* scan_tmp, hipStream_t s,
                          u64** result);
hipError_t launch_flag_unique(const u64* k, u64 n, u64 sentinel, uint32_t* flags, hipStream_t s);
hipError_t launch_flag_absent(const u64* b, u64 nb, const u64* a, uint32_t na, uint32_t* flags, hipStream_t s);
hipError_t launch_compact(const u64* k, u64 n, const uint32_t* flags, const uint32_t* pos, u64* out, hipStream_t s);
hipError_t launch_parent_keys(const u64* c, u64 n, const BitDomain& dc, int ld, const BitDomain& dp, u64* out,
                              hipStream_t s);
struct KeyLevel {
    BitDomain dom;
    const u64* A;    // sorted keys of the own-brick cells
    uint32_t na;
    const u64* E;    // sorted keys of the dilation-only cells
    uint32_t ne;
    int lev;
    int range;
};
hipError_t launch_emit_list(const KeyLevel& L, Node* nodes, int has_clist, hipStream_t s);
hipError_t launch_link_list(const KeyLevel& C, const KeyLevel& P, int parent_is_root, int ld, Node* cnodes,
                            u64* pclists, u64 clist_len, hipStream_t s);

struct RootArgs {
    Node* root;
    int pos[3];
    int lev;
};
hipError_t launch_write_root_multi(const RootArgs* a, int G, Arena* ar, hipStream_t s);

// ---- incremental topology (ActivateIncreBricksGPU, topology.hip)
// Per point w = p + o >= 0 and level L < D, the nodes covering
// [int(w) - r, int(w) + r]: missing ones set their bit in bm[L] (domain
// dom[L], node coordinates), found leaves get marker 1.
struct IncreMarkArgs {
    Tree t;
    PointSrc pts;
    float o[3];
    float radius;
    int D;
    BitDomain dom[MAXLEV];
    uint32_t* bm[MAXLEV];
    uint32_t* marker;
};
hipError_t launch_incre_mark(const IncreMarkArgs& a, hipStream_t s);
hipError_t launch_apply_markers(Node* leaves, const uint32_t* marker, u64 n, hipStream_t s);
// new nodes of one level from its bitmap: node base + rank, child list cbase + rank
struct IncreEmitArgs {
    BitDomain dom;
    const uint32_t* bm;
    const uint32_t* rank;
    Node* nodes;
    int lev, range, has_clist;
    u64 base, cbase;
};
hipError_t launch_incre_emit(const IncreEmitArgs& a, hipStream_t s);
// the incremental build over keys (keysort.hip): missing cells as
// (level << 56 | cell) keys, counted (and leaf markers set) then written
hipError_t launch_incre_keys(const IncreMarkArgs& a, uint32_t* cnt, const uint32_t* off, u64* keys, bool emit,
                             hipStream_t s);
hipError_t launch_level_counts(const u64* k, u64 n, uint32_t* counts, hipStream_t s);
hipError_t launch_incre_emit_list(const u64* k, uint32_t n, const BitDomain& d, Node* nodes, int lev, int range,
                                  int has_clist, u64 base, u64 cbase, hipStream_t s);
// gvdbLinkBricks / gvdbDelinkLeafBricks / gvdbDelinkBricks with tree walks
hipError_t launch_link_walk(const Tree& t, int lev, hipStream_t s);
hipError_t launch_delink_leaves(const Tree& t, hipStream_t s);
hipError_t launch_delink_level(const Tree& t, int lev, hipStream_t s);
// *out += leaves with mFlags == 0 (out must start at zero)
hipError_t launch_count_unflagged(const Tree& t, uint32_t* out, hipStream_t s);

// ---- Compute operators (compute.hip): effect ids of the reference
// (src/gvdb_volume_gvdb.h:201-208) for float/half channels
enum { OP_FILL = 110, OP_SMOOTH = 113, OP_NOISE = 114, OP_GROW = 115 };
struct ComputeArgs {
    Tree t;               // atlas geometry (brick_res, atlas_cnt)
    void* atlas;
    int atlas_type;
    int op;
    float p1, p2, p3;
    int slots;            // smooth/noise/grow: atlas slots 0 .. slots-1
    int lo[3], hi[3];     // fill: inclusive atlas-voxel range
};
hipError_t launch_compute(const ComputeArgs& a, hipStream_t s);

// ---- atlas & points (points.hip)
struct AtlasArgs {
    Tree t;
    const uint32_t* slot_scan;
    int n;
};
hipError_t launch_update_atlas_multi(const AtlasArgs* a, int G, Arena* ar, hipStream_t s);

// flags[i] = leaf i in use (mode 0: mFlags) / linked (mode 1: mParent != ID_UNDEF64)
struct FlagArgs {
    Tree t;
    uint32_t* flags;
    int n;
    int mode;
};
hipError_t launch_leaf_flags_multi(const FlagArgs* a, int G, Arena* ar, hipStream_t s);

// ComputeBounds over flagged leaves: out6 = min corner, max corner + range
struct BoundsArgs {
    Tree t;
    int n;
    int* out6;
};
hipError_t launch_leaf_bounds_multi(const BoundsArgs* a, int G, Arena* ar, hipStream_t s);

// Which points a subcell insert bins (w = p + trans):
//   BIN_FLOAT (InsertPointsSubcell): the reference's count pass skips
//     p.z == NOHIT and w < 0, its insert pass p < 0; binned when both agree;
//   BIN_FP16 (InsertPointsSubcell_FP16): the count pass as above, the insert
//     pass skips w.z == NOHIT and w < 0; binned when both agree.
enum { BIN_FLOAT = 0, BIN_FP16 = 1 };

// push binning (the reference's two atomic passes), one grid per launch
hipError_t launch_count_subcell(const Tree& t, PointSrc pts, int scs, int spb, const float trans[3], float radius,
                                const uint32_t* mapping, uint32_t* sc_cnt, int rule, hipStream_t s);
hipError_t launch_insert_subcell(const Tree& t, PointSrc pts, int scs, int spb, const float trans[3], float radius,
                                 const uint32_t* mapping, uint32_t* sc_cnt, const uint32_t* sc_off,
                                 float* sc_pnt, int rule, hipStream_t s);

// InsertPointsSubcell_FP16's entry format (gvdbInsertSubcell_fp16,
// kernels/cuda_gvdb_particles.cuh:154-216): each coordinate as
// u16((w - pos_min) / pos_range * 65535), here saturated to [0, 65535];
// n3 = 3 x entries.
hipError_t launch_quantize_entries(const float* src, uint16_t* dst, u64 n3, const float qmin[3], const float qrange[3],
                                   hipStream_t s);

struct SubPosArgs {
    Tree t;
    int scs, spb;
    const uint32_t* mapping;
    int* pos;
    int* nid;
    int n;
};
hipError_t launch_subcell_pos_multi(const SubPosArgs* a, int G, Arena* ar, hipStream_t s);

// sc_use[s]: entries at the head of subcell s that can raise a voxel (all of
// them when the push path binned the points); accumulate mode reads sc_cnt.
struct GatherArgs {
    Tree t;
    int scs;
    float radius;
    const int* sc_nid;
    const uint32_t* sc_cnt;
    const uint32_t* sc_use;
    const uint32_t* sc_off;
    const int* sc_pos;
    const float* sc_pnt;
    void* atlas;
    int atlas_type;
    int accumulate;
    int nbricks;          // linked bricks (one workgroup each)
    int levelset;         // 1: GatherLevelSet, min(3, min (|p - v| - r)) over every entry
    // GatherLevelSet_FP16: u16 entries, position = u / 65535 * qrange + qmin
    // (sc_pnt unused)
    const uint16_t* sc_pnt16;
    float qmin[3], qrange[3];
};
hipError_t launch_gather_multi(const GatherArgs* a, int G, Arena* ar, hipStream_t s);

// pull-based binning: bucket points by their own subcell, then every brick
// pulls from the subcells around it
struct BucketArgs {
    Tree t;
    PointSrc pts;
    int scs, spb;
    float tr[3];
    uint32_t* key;
    uint32_t* slot;
    uint32_t* bcnt;
    uint32_t* orphans;
    // optional: adds the number of (point, subcell) entries InsertSubcell
    // would make (u64 counter); needs the neighbour masks and R <= spd
    const uint32_t* nmask;
    float radius;
    u64* len;             // 64 partial counters, summed by launch_sum_parts_multi
    int rule;             // BIN_FLOAT or BIN_FP16
    BrickTab bt;          // leaf lookup (tab == nullptr: the walk)
};
hipError_t launch_sum_parts_multi(const u64* const* parts, u64* const* out, int G, Arena* ar, hipStream_t s);
hipError_t launch_bucket_points_multi(const BucketArgs* a, int G, Arena* ar, hipStream_t s);

struct ScatterArgs {
    PointSrc pts;
    PointSrc clr;            // optional (base null): a uint32 colour per point, kept in the fourth word
    float tr[3];
    const uint32_t* key;
    const uint32_t* slot;
    const uint32_t* boff;
    float4* sorted;          // bucketed positions (+trans), 16 bytes each
};
hipError_t launch_bucket_scatter_multi(const ScatterArgs* a, int G, Arena* ar, hipStream_t s);

int pull_cells(int sd, int R);
int pull_max_cells();
int pull_max_spb();
struct PullArgs {
    Tree t;
    int nleaves;
    int scs, R;
    float radius;
    const int* nbr;
    const uint32_t* bcnt;
    const uint32_t* boff;
    const float4* sorted;
    const uint32_t* mapping;
    uint32_t* sc_cnt;
    uint32_t* sc_use;
    const uint32_t* sc_off;
    float* sc_pnt;
};
hipError_t launch_pull_multi(int write, const PullArgs* a, int G, Arena* ar, hipStream_t s);

// GatherDensity straight from the buckets: one 512-thread workgroup per
// linked 8^3 brick with 4^3 subcells, reading the points of the subcells
// around it (the same candidates the pull binning reads).
struct BucketGatherArgs {
    Tree t;
    int nleaves;
    int R;
    float radius;
    const int* nbr;
    const uint32_t* bcnt;
    const uint32_t* boff;
    const float4* sorted;
    void* atlas;
    int atlas_type;
    int accumulate;
    void* faces;          // optional (scatter kernel): the six 8x8 boundary faces of each brick, channel type
    int levelset;         // 1: GatherLevelSet (nearest-entry kernel)
};
hipError_t launch_bucket_gather_multi(const BucketGatherArgs* a, int G, Arena* ar, hipStream_t s);

// GatherDensity / GatherLevelSet colour channel (chanClr) from the colours
// the bucketed insert carried in each point's fourth word: every voxel of a
// subcell gets the mean of its entry list's colours (uchar4 per voxel).
struct ColorArgs {
    Tree t;
    int nleaves;
    int R;
    float radius;
    int scs;
    const int* nbr;
    const uint32_t* bcnt;
    const uint32_t* boff;
    const float4* sorted;
    void* atlas;          // the uchar4 channel
    int accumulate;
};
hipError_t launch_gather_color_multi(const ColorArgs* a, int G, Arena* ar, hipStream_t s);
// Face buffer layout: brick b's faces at b * FACE_VALS; face f = x0, x7, y0,
// y7, z0, z7 (local interior coordinate 0 or 7) at f * 64, each face row-major
// in its other two coordinates, the higher one first (x faces: z*8 + y; y
// faces: z*8 + x; z faces: y*8 + x).
static const int FACE_VALS = 384;

// ---- apron (apron.hip)
// nbr[b*27 + d]: leaf holding the neighbour of brick b in direction
// d = (dz+1)*9 + (dy+1)*3 + (dx+1), or -1; d = 13 is b itself, or -1 when the
// walk from the top node does not reach b
struct NbrArgs {
    Tree t;
    int* nbr;
    int n;
    uint32_t* mask;       // optional: bit d set when nbr[b*27 + d] >= 0 (written whole)
    long long* base;      // optional: atlas offset of each neighbour's brick (d = 13: own), -1 if none
    BrickTab bt;          // leaf lookup (tab == nullptr: the walks)
};
hipError_t launch_neighbors_multi(const NbrArgs* a, int G, Arena* ar, hipStream_t s);

struct ApronArgs {
    Tree t;
    const long long* base;   // neighbour brick offsets from the neighbour pass
    const int* nbr;          // neighbour leaves (read with faces)
    const void* faces;       // optional: the gather's face buffer, current with the atlas
    void* atlas;
    int atlas_type;
    float boundval;
    int n;
};
hipError_t launch_apron_multi(const ApronArgs* a, int G, Arena* ar, hipStream_t s);

// UpdateApron without allocation: the neighbour table (27 atlas offsets of
// 8 bytes, then 27 leaf indices of 4 bytes, per brick) lives in a fixed
// device array of NOALLOC_SCRATCH_BYTES per device instead of memory the
// library allocates, so at most NOALLOC_MAX_BRICKS bricks fit.
static const size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
static const int NOALLOC_BYTES_PER_BRICK = 27 * (8 + 4);
static const int NOALLOC_MAX_BRICKS = int(NOALLOC_SCRATCH_BYTES / NOALLOC_BYTES_PER_BRICK);
// The device address of `device`'s copy of the array, looked up once per
// device (with that device current) and kept; later calls return the kept
// address.
hipError_t noalloc_scratch(int device, void** ptr);
// The neighbour pass and the apron copy for one grid of `bricks` bricks,
// with the neighbour table in `scratch` (the address noalloc_scratch gave);
// `faces` as in ApronArgs.
hipError_t launch_apron_noalloc(const Tree& t, void* scratch, const void* faces, void* atlas, int atlas_type,
                                float boundval, int bricks, hipStream_t s);

// ---- render (render.hip)
struct RenderArgs {
    Tree t;
    ScnBlock s;
    const float4* tf;
    const void* atlas;
    void* out;
    int out_float;
    int shading;
    int atlas_type;
    int tile_w;          // pixel tile of one 64-lane workgroup: tile_w x (64 / tile_w); set by the launcher
    const uint32_t* clr; // the colour channel (uchar4 per voxel, SetColorChannel), or null
};
hipError_t launch_render_multi(const RenderArgs* a, int G, Arena* ar, hipStream_t s);

// ---- launch-shape knobs (tune.cpp), process-wide, picked by the autotuner
enum Knob {
    KNOB_RENDER_UNROLL_EA = 0,   // samples fetched ahead per brick step, emission-absorption: 1, 2, 4
    KNOB_RENDER_UNROLL_ISO,      // the same for the isosurface march
    KNOB_RENDER_TILE_EA,         // ray-packet width of a 64-lane workgroup (8x8, 16x4, 32x2, 64x1): 8, 16, 32, 64
    KNOB_RENDER_TILE_ISO,
    KNOB_GATHER_THREADS,         // threads per brick of the scatter gather at radius <= 3: 128, 256
    KNOB_APRON_RUN,              // bricks per workgroup of the multi-grid apron: 1, 2, 4, 8
    KNOB_XCD_ORDER,              // per-brick kernels: 0 launch order, 1 each XCD a contiguous range of bricks
    KNOB_APRON_QUEUE,            // multi-grid apron: 0 static runs, 1 persistent workgroups over a brick queue
    KNOB_MARK_PATH,              // topology marks: 0 by domain size, 1 a byte per brick, 2 direction masks, 3 bits,
                                 // 4 sorted brick keys (the build taken above 2^34 bricks)
    KNOB_APRON_RUN_ONE,          // bricks per workgroup of the one-grid apron: 1, 2, 4, 8
    KNOB_COUNT
};
int knob(int k);
bool set_knob(int k, int v);          // false (and unchanged) when v is not one of the knob's values
static const int KNOB_MAX_VALUES = 8;
int knob_values(int k, int out[KNOB_MAX_VALUES]);   // the knob's values; returns how many
class Volume;
class Batch;
// The autotuner (tune.cpp): times the caller's own workload per knob value
// and keeps the fastest.  Render: tile x unroll (ms12 in that order, best2 =
// {unroll, tile width}); unroll alone; gather: threads per brick; batched
// stages 0 gather, 1 apron, 2 render.
int tune_render(Volume& v, int shading, int chan, int buf, int* best2, float* ms12);
int tune_render_unroll(Volume& v, int shading, int chan, int buf, int* best, float* ms3);
int tune_gather(Volume& v, int scs, int n, float radius, const float trans[3], int len, int chan, int* best,
                float* ms2);
int tune_batch(Batch& B, Volume* const* vs, int G, int stage, int scs, const int* len, const float* radius, int chan,
               float boundval, const int* shading, const int* chans, const int* bufs, int* best, float* ms,
               int* failed);

}  // namespace gvdb
