// tree_device.h -- GVDB tree walks on the device.
//
// These follow the reference's walks exactly where its behaviour is defined
// (kernels/cuda_gvdb_nodes.cuh:115-124 getChild, :199-253 getNode /
// getNodeAtPoint, :255-288 getPosLeafParent).  Where the reference would
// index a child list outside [0, res) (positions outside the start node), the
// walks here report "no node" instead of reading neighbouring memory.
#pragma once
#include "gvdb_device.h"

namespace gvdb {

__device__ inline Node* node_at(const Tree& t, int lev, u64 idx) { return t.nodes[lev] + idx; }

// Returns the child index, -1 for an empty slot or a node without a list.
__device__ inline int child_of(const Tree& t, const Node* n, int b) {
    u64 listid = n->mChildList;
    if (listid == ID_UNDEFL) return -1;
    int clev = int((listid >> 8) & 0xFF);
    u64 cndx = listid >> 16;
    u64 r = u64(t.res[clev]);
    u64 v = t.clists[clev][cndx * r * r * r + u64(b)];
    return int(v >> 16);
}

// Leaf index of the brick containing integer position `pos`, or -1.
// Node ranges are powers of two, so (d * res) / range (C division, towards
// zero) is d shifted right by log2(range / res), with the sign applied after.
__device__ inline int child_cell(long long d, int sh) { return d >= 0 ? int(d >> sh) : -int((-d) >> sh); }

__device__ inline int pos_leaf(const Tree& t, int px, int py, int pz) {
    const Node* nd = node_at(t, t.top_lev, 0);
    int l = nd->mLev;
    if (l == 0) return 0;
    while (true) {
        int res = t.res[l];
        const int sh = __builtin_ctz(unsigned(t.range[l])) - t.logdim[l];
        int a = child_cell((long long)px - nd->mPos[0], sh);
        int b = child_cell((long long)py - nd->mPos[1], sh);
        int c = child_cell((long long)pz - nd->mPos[2], sh);
        if (a < 0 || b < 0 || c < 0 || a >= res || b >= res || c >= res) return -1;
        int child = child_of(t, nd, (c * res + b) * res + a);
        if (child < 0) return -1;
        l--;
        if (l == 0) return child;
        nd = node_at(t, l, u64(child));
    }
}

// The leaf at brick coordinates (bx, by, bz) from a full rebuild's brick
// table, or -1 outside its box or where no leaf is.  Within that topology it
// equals pos_leaf at the brick's corner and point_leaf inside it: every leaf
// is reachable, and the box holds every leaf.
__device__ inline u64 tab_index(const BrickTab& b, unsigned x, unsigned y, unsigned z) {
    return (u64(x) * unsigned(b.dim[1]) + y) * unsigned(b.dim[2]) + z;
}
__device__ inline int tab_leaf(const BrickTab& b, int bx, int by, int bz) {
    const unsigned x = unsigned(bx - b.lo[0]), y = unsigned(by - b.lo[1]), z = unsigned(bz - b.lo[2]);
    if (x >= unsigned(b.dim[0]) || y >= unsigned(b.dim[1]) || z >= unsigned(b.dim[2])) return -1;
    const uint32_t v = uint32_t(b.tab[tab_index(b, x, y, z)]);
    return v == ~0u ? -1 : int(v);
}

// getPosParent (kernels/cuda_gvdb_nodes.cuh:303-348): the node at stopLev
// on the walk from the top node to `pos`, or -1.  A stopLev at or above the
// top node's level is never reached, as in the reference.
__device__ inline int pos_parent(const Tree& t, int px, int py, int pz, int stopLev) {
    const Node* nd = node_at(t, t.top_lev, 0);
    int l = nd->mLev;
    if (l == 0) return 0;
    while (l > 0) {
        int res = t.res[l];
        const int sh = __builtin_ctz(unsigned(t.range[l])) - t.logdim[l];
        int a = child_cell((long long)px - nd->mPos[0], sh);
        int b = child_cell((long long)py - nd->mPos[1], sh);
        int c = child_cell((long long)pz - nd->mPos[2], sh);
        if (a < 0 || b < 0 || c < 0 || a >= res || b >= res || c >= res) return -1;
        int child = child_of(t, nd, (c * res + b) * res + a);
        l--;
        if (child < 0) return -1;
        if (l == stopLev) return child;
        nd = node_at(t, l, u64(child));
    }
    return -1;
}

// getParent (kernels/cuda_gvdb_nodes.cuh:350-397): the node at stopLev
// holding `pos`; the top node itself (0) when stopLev is its level or it is
// unflagged; -1 when the walk meets an empty slot (where the reference reads
// past it) or leaves the node.
__device__ inline int get_parent(const Tree& t, int stopLev, const int pos[3]) {
    const Node* nd = node_at(t, t.top_lev, 0);
    if (!nd->mFlags) return 0;
    int l = nd->mLev;
    if (l == stopLev) return 0;
    while (l > stopLev) {
        int res = t.res[l];
        const int sh = __builtin_ctz(unsigned(t.range[l])) - t.logdim[l];
        int a = child_cell((long long)pos[0] - nd->mPos[0], sh);
        int b = child_cell((long long)pos[1] - nd->mPos[1], sh);
        int c = child_cell((long long)pos[2] - nd->mPos[2], sh);
        if (a < 0 || b < 0 || c < 0 || a >= res || b >= res || c >= res) return -1;
        l--;
        int child = child_of(t, nd, (c * res + b) * res + a);
        if (child < 0) return -1;
        if (l == stopLev) return child;
        nd = node_at(t, l, u64(child));
    }
    return -1;
}

// getNodeAtPoint: float walk with per-level bounds, returns the leaf or -1.
__device__ inline int point_leaf(const Tree& t, float px, float py, float pz) {
    int lev = t.top_lev;
    const Node* nd = node_at(t, lev, 0);
    int id = 0;
    while (lev > 0) {
        float vx = float(nd->mPos[0]), vy = float(nd->mPos[1]), vz = float(nd->mPos[2]);
        float rg = float(t.range[lev]);
        if (px < vx || py < vy || pz < vz || px >= vx + rg || py >= vy + rg || pz >= vz + rg) return -1;
        // division by the child range, a power of two: the same as a
        // multiplication by its exact reciprocal
        const float inv = __builtin_ldexpf(1.0f, t.logdim[lev] - __builtin_ctz(unsigned(t.range[lev])));
        int a = int((px - vx) * inv), b = int((py - vy) * inv), c = int((pz - vz) * inv);
        int bit = (((c << t.logdim[lev]) + b) << t.logdim[lev]) + a;
        lev--;
        int child = child_of(t, nd, bit);
        if (child < 0) return -1;
        id = child;
        nd = node_at(t, lev, u64(child));
    }
    return id;
}

// Atlas slot of a node whose interior corner is mValue.
__device__ inline u64 slot_of(const Tree& t, const int mv[3]) {
    int br = t.brick_res;
    return (u64(mv[2] / br) * t.atlas_cnt[1] + u64(mv[1] / br)) * t.atlas_cnt[0] + u64(mv[0] / br);
}

}  // namespace gvdb
