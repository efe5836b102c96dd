// kernels.h -- host-side launchers for the gfx950 kernels of the point path.
// Every launcher enqueues on the given stream and returns the launch status.
#pragma once
#include <hip/hip_runtime.h>
#include "gvdb_device.h"

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
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, u64 n, hipStream_t s);
hipError_t launch_fill_u64(u64* p, u64 v, u64 n, hipStream_t s);
// exclusive scan of u32 values; `mode` selects what is scanned:
//   0: in0[i];  1: popcount(in0[i]);  2: popcount(in1[i] & ~in0[i]);  3: in0[i] != 0
// `total` (device, may be null) receives the sum.  `tmp` must hold
// scan_tmp_words(n) words.
u64 scan_tmp_words(u64 n);
hipError_t launch_scan_u32(const uint32_t* in0, const uint32_t* in1, int mode, uint32_t* out, u64 n,
                           uint32_t* total, uint32_t* tmp, hipStream_t s);

// ---- topology (topology.hip)
hipError_t launch_bbox(PointSrc pts, const float orig[3], uint32_t* keys6, hipStream_t s);
// masks: optional zeroed u32 per level-0 domain brick (nbits words)
hipError_t launch_mark_points(PointSrc pts, const float orig[3], float radius, int dilate, int range0,
                              BitDomain d0, uint32_t* bmA, uint32_t* bmB, uint32_t* masks, hipStream_t s);
hipError_t launch_mark_parents(BitDomain dc, const uint32_t* bmA_c, const uint32_t* bmB_c, int shift, BitDomain dp,
                               uint32_t* bmA_p, uint32_t* bmB_p, hipStream_t s);
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
hipError_t launch_emit_nodes(const LevelBuild& L, Node* nodes, int has_clist, hipStream_t s);
hipError_t launch_link_level(const LevelBuild& child, const LevelBuild* parent, int parent_is_root,
                             int logdim_parent, Node* cnodes, u64* parent_clists, u64 clist_len, hipStream_t s);
hipError_t launch_write_root(Node* root, const int pos[3], int lev, hipStream_t s);

// ---- atlas & points (points.hip)
hipError_t launch_update_atlas(Tree t, const uint32_t* slot_scan, int nleaves, hipStream_t s);
// flags[i] = (leaf i linked: mParent != ID_UNDEF64) / (leaf i in use: mFlags != 0)
hipError_t launch_leaf_flags(Tree t, uint32_t* flags, int nleaves, hipStream_t s);
hipError_t launch_leaf_used(Tree t, uint32_t* flags, int nleaves, hipStream_t s);
// ComputeBounds over flagged leaves: out6 = min corner, max corner + range
hipError_t launch_leaf_bounds(Tree t, int nleaves, int* out6, hipStream_t s);
hipError_t launch_count_subcell(Tree t, PointSrc pts, int scs, int spb, const float trans[3], float radius,
                                const uint32_t* mapping, uint32_t* sc_cnt, hipStream_t s);
hipError_t launch_insert_subcell(Tree t, PointSrc pts, int scs, int spb, const float trans[3], float radius,
                                 const uint32_t* mapping, uint32_t* sc_cnt, const uint32_t* sc_off,
                                 float* sc_pnt, hipStream_t s);
hipError_t launch_subcell_pos(Tree t, int scs, int spb, const uint32_t* mapping, int* sc_pos, int* sc_nid,
                              int nleaves, hipStream_t s);
// sc_use[s]: entries at the head of subcell s that can raise a voxel (all of
// them when the push path binned the points); accumulate mode reads sc_cnt.
hipError_t launch_gather_density(Tree t, int scs, int nbricks_sc, float radius, const int* sc_nid,
                                 const uint32_t* sc_cnt, const uint32_t* sc_use, const uint32_t* sc_off,
                                 const int* sc_pos, const float* sc_pnt, void* atlas, int atlas_type, int accumulate,
                                 hipStream_t s);

// pull-based binning (points.hip): bucket points by their own subcell, then
// every brick pulls from the subcells around it
hipError_t launch_bucket_points(Tree t, PointSrc pts, int scs, int spb, const float trans[3], uint32_t* key,
                                uint32_t* slot, uint32_t* bcnt, uint32_t* orphans, hipStream_t s);
hipError_t launch_bucket_scatter(PointSrc pts, const float trans[3], const uint32_t* key, const uint32_t* slot,
                                 const uint32_t* boff, float* sorted, hipStream_t s);
int pull_cells(int sd, int R);
int pull_max_cells();
int pull_max_spb();
hipError_t launch_pull(int write, Tree t, int nleaves, int scs, int R, float radius, const int* nbr27,
                       const uint32_t* bcnt, const uint32_t* boff, const float* sorted, const uint32_t* mapping,
                       uint32_t* sc_cnt, uint32_t* sc_use, const uint32_t* sc_off, float* sc_pnt, hipStream_t s);

// ---- apron (apron.hip)
// nbr27[b*27 + d]: leaf holding the neighbour of brick b in direction
// d = (dz+1)*9 + (dy+1)*3 + (dx+1), or -1; d = 13 is b itself, or -1 when the
// walk from the top node does not reach b
hipError_t launch_neighbors(Tree t, int* nbr27, int nbricks, hipStream_t s);
hipError_t launch_update_apron(Tree t, const int* nbr27, int nbricks, void* atlas, int atlas_type, float boundval,
                               hipStream_t s);

// ---- render (render.hip)
hipError_t launch_render(Tree t, const ScnBlock& scn, const float4* tf, const void* atlas, int atlas_type,
                         int shading, void* out, int out_float, hipStream_t s);

}  // namespace gvdb
