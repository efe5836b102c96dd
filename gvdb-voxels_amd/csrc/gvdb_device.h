// gvdb_device.h -- data layout shared by the host runtime and the gfx950 kernels.
//
// Node pools and the brick atlas keep GVDB's layout (src/gvdb_node.h:27-40,
// src/gvdb_allocator.h:59-62, kernels/cuda_gvdb_nodes.cuh:24-67 of the
// reference): 64-byte nodes in one pool per level (group 0), dense child
// lists of res^3 u64 references per interior node (group 1), references
// packed as grp | lev<<8 | idx<<16.  The atlas keeps GVDB's logical
// coordinates (brick slot id -> (x,y,z) in a cnt.x * cnt.y * cnt.z grid of
// padded bricks, node mValue = slot corner + apron); physically each padded
// brick is stored contiguously (brick-major), see atlas_offset().
#pragma once
#include <stdint.h>

#ifndef __HIPCC__
#define __host__
#define __device__
#endif

namespace gvdb {

typedef unsigned long long u64;

static const u64 ID_UNDEFL = 0xFFFFFFFFull;
static const u64 ID_UNDEF64 = 0xFFFFFFFFFFFFFFFFull;
static const int MAXLEV = 10;
static const int MAX_ITER = 256;
static const int TF_SIZE = 16384;
static const float NOHIT = 1.0e10f;

// shading modes (kernels/cuda_gvdb_scene.cuh:25-32 of the reference)
enum { SHADE_VOXEL = 0, SHADE_TRILINEAR = 4, SHADE_VOLUME = 7 };
// channel types (src/gvdb_types.h of the reference)
enum { T_UCHAR = 0, T_UCHAR4 = 2, T_FLOAT = 3, T_HALF = 100 };
static const int CHAN_UNDEF = 255;   // no colour channel (kernels/cuda_gvdb_nodes.cuh:21)

struct Node {
    uint8_t mLev, mFlags, mPriority, pad;
    int32_t mPos[3];
    int32_t mValue[3];
    float mVRange[3];
    u64 mParent;
    u64 mChildList;
    u64 mMask;
};
static_assert(sizeof(Node) == 64, "GVDB node is 64 bytes");

struct AtlasNode {
    int32_t mPos[3];
    int32_t mLeafNode;
};

__host__ __device__ inline u64 Elem(int grp, int lev, u64 ndx) { return u64(grp) | (u64(lev) << 8) | (ndx << 16); }

// The leaf index of every brick of the level-0 box of a full rebuild, in the
// box's bit order (index ((x - lo[0]) * dim[1] + (y - lo[1])) * dim[2] +
// (z - lo[2]), brick coordinates), ~0u where no leaf is: written by the node
// emission, so a lookup replaces the walk from the top node while the
// topology is that build's (every leaf of it is reachable from the top
// node).  The high word of an entry holds the leaf's 27-bit neighbour mask
// once the neighbour pass has run on that topology (0 before).  tab ==
// nullptr: walk.
struct BrickTab {
    u64* tab;   // written by the emission and the neighbour pass, read by lookups
    int lo[3];
    int dim[3];
};
static const u64 BRICK_TAB_MAX = u64(1) << 24;   // box bricks with a table (the per-bit emission)

// Device view of one grid (the role of the reference's VDBInfo).  Passed to
// kernels by value.
struct Tree {
    int levels;
    int top_lev;
    int logdim[MAXLEV];
    int res[MAXLEV];
    int range[MAXLEV];
    int nodecnt[MAXLEV];
    Node* nodes[MAXLEV];       // group 0, one pool per level
    u64* clists[MAXLEV];       // group 1, res^3 entries per list
    // atlas (channel data is passed separately)
    AtlasNode* amap;
    int atlas_cnt[3];
    int brick_res;             // padded brick side (brickwid + 2 apron)
    int brickwid;
    int apron;
    int brick_vox;             // brick_res^3
    float bmin[3], bmax[3];
    float epsilon;
};

// Brick-major atlas addressing: voxel (ax,ay,az) in atlas coordinates.
__host__ __device__ inline u64 atlas_offset(const Tree& t, int ax, int ay, int az) {
    int br = t.brick_res;
    int sx = ax / br, sy = ay / br, sz = az / br;
    int lx = ax - sx * br, ly = ay - sy * br, lz = az - sz * br;
    u64 slot = (u64(sz) * t.atlas_cnt[1] + sy) * t.atlas_cnt[0] + sx;
    return slot * u64(t.brick_vox) + u64((lz * br + ly) * br + lx);
}

// Scene block (PrepareRender, src/gvdb_volume_gvdb.cpp:4254-4306 of the
// reference).  68 32-bit words; the layout is part of the C ABI.
struct ScnBlock {
    int32_t width, height;
    float campos[3], cams[3], camu[3], camv[3];
    float light_pos[3];
    float backclr[4];
    float extinct[3], steps[3], cutoff[3], thresh[3], shadow[3];
    float invxform[16], invxrot[16];
};
static_assert(sizeof(ScnBlock) == 68 * 4, "scene block is 68 words");

}  // namespace gvdb
