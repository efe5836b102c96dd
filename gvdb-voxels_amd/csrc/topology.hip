// topology.hip -- dynamic topology on gfx950.
//
// The reference (RebuildTopology, src/gvdb_volume_gvdb.cpp:1181-1413 and
// ActivateExtraBricksGPU :931-1016) builds per-point u64 brick keys for every
// level, sorts them with Thrust, marks and compacts the unique keys, copies
// them to the host, allocates one node per key in sorted order on the host,
// commits the pools and links each level with a root-to-node walk.
//
// Here the key set of each level is a bitmap over the box of node coordinates
// the points can reach, stored x-major so that bit order is exactly GVDB's
// sorted key order (lev, x, y, z).  Setting a bit replaces sort+unique, a
// popcount scan of the words gives every node its pool index (its rank), and
// parents are found by a rank query instead of a tree walk.  The dilation
// pass of the reference (keys of nodes within [int(p)-r, int(p)+2r] that are
// not in the first pass, appended after it in their own sorted order) is a
// second bitmap per level whose ranks start after the first pass's count.
// Higher levels are marked from the set bits of the level below, which is the
// same set because power-of-two node ranges nest.
#include "kernels.h"

namespace gvdb {

__device__ inline uint32_t float_key(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ inline const float* point_at(const PointSrc& p, int i) {
    return reinterpret_cast<const float*>(p.base + u64(i) * p.stride + p.off);
}

// GetBoundingBox's split + Thrust min/max (src/gvdb_volume_gvdb.cpp:5404-5467).
__global__ __launch_bounds__(256) void k_bbox(PointSrc pts, float ox, float oy, float oz, uint32_t* keys) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < pts.n; i += gridDim.x * blockDim.x) {
        const float* p = point_at(pts, i);
        float w[3] = {p[0] + ox, p[1] + oy, p[2] + oz};
#pragma unroll
        for (int a = 0; a < 3; a++) { mn[a] = fminf(mn[a], w[a]); mx[a] = fmaxf(mx[a], w[a]); }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], d, 64));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], d, 64));
        }
    }
    __shared__ float smn[4][3], smx[4][3];
    int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; a++) { smn[wave][a] = mn[a]; smx[wave][a] = mx[a]; }
    __syncthreads();
    if (threadIdx.x < 3) {
        int a = threadIdx.x;
        float m0 = smn[0][a], m1 = smx[0][a];
        for (int w = 1; w < 4; w++) { m0 = fminf(m0, smn[w][a]); m1 = fmaxf(m1, smx[w][a]); }
        atomicMin(&keys[a], float_key(m0));
        atomicMax(&keys[3 + a], float_key(m1));
    }
}

hipError_t launch_bbox(PointSrc pts, const float orig[3], uint32_t* keys6, hipStream_t s) {
    hipError_t e = launch_fill_u32(keys6, 0xFFFFFFFFu, 3, s);
    if (e != hipSuccess) return e;
    e = launch_fill_u32(keys6 + 3, 0u, 3, s);
    if (e != hipSuccess) return e;
    int blocks = (pts.n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_bbox, dim3(blocks), dim3(256), 0, s, pts, orig[0], orig[1], orig[2], keys6);
    return hipGetLastError();
}

__device__ inline bool dom_index(const BitDomain& d, int x, int y, int z, u64& idx) {
    int a = x - d.lo[0], b = y - d.lo[1], c = z - d.lo[2];
    if (a < 0 || b < 0 || c < 0 || a >= d.dim[0] || b >= d.dim[1] || c >= d.dim[2]) return false;
    idx = (u64(a) * d.dim[1] + b) * d.dim[2] + c;
    return true;
}

__device__ inline void dom_coords(const BitDomain& d, u64 idx, int& x, int& y, int& z) {
    u64 yz = u64(d.dim[1]) * d.dim[2];
    u64 a = idx / yz;
    u64 r = idx - a * yz;
    u64 b = r / d.dim[2];
    x = int(a) + d.lo[0];
    y = int(b) + d.lo[1];
    z = int(r - b * d.dim[2]) + d.lo[2];
}

__device__ inline void set_bit(uint32_t* bm, u64 idx) {
    uint32_t m = 1u << (idx & 31);
    uint32_t* w = bm + (idx >> 5);
    if (!(__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & m))
        __hip_atomic_fetch_or(w, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// gvdbCalcBrickId (kernels/cuda_gvdb_particles.cuh:352-382) for level 0 and
// gvdbCalcExtraBrickId (:467-512) for the dilation.  Each point contributes
// its brick and, for the dilation, the bricks covering [int(p)-r, int(p)+2r]
// on every axis.  When those lie within one brick of the point's own (r up to
// one brick), the point ORs a 27-bit direction mask into its own brick's word
// (`masks`, bit 13 = the brick itself), which k_expand_masks turns into bitmap
// bits; this replaces up to 8 bitmap atomics per point by one rarely taken
// atomic.  Points with wider reach (or no mask array) set the bits directly.
__global__ __launch_bounds__(256) void k_mark_points(PointSrc pts, float ox, float oy, float oz, float radius,
                                                     int dilate, int range0, float inv_range0, BitDomain d0,
                                                     uint32_t* bmA, uint32_t* bmB, uint32_t* masks) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < pts.n; i += gridDim.x * blockDim.x) {
        const float* p = point_at(pts, i);
        float w[3] = {p[0] + ox, p[1] + oy, p[2] + oz};
        if (!isfinite(w[0]) || !isfinite(w[1]) || !isfinite(w[2])) continue;
        if (w[0] < 0 || w[1] < 0 || w[2] < 0) continue;
        int c[3];
#pragma unroll
        for (int a = 0; a < 3; a++) c[a] = int(floorf(w[a] * inv_range0));
        u64 idx;
        const bool inside = dom_index(d0, c[0], c[1], c[2], idx);
        int lo[3] = {c[0], c[1], c[2]}, hi[3] = {c[0], c[1], c[2]};
        if (dilate) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                int iw = int(w[a]);
                int bmin = int(float(iw) - radius);
                int bmax = int(float(iw) + 2.0f * radius);
                lo[a] = bmin / range0;   // C truncation, as the reference's int division
                hi[a] = bmax / range0;
            }
        }
        const bool near = lo[0] >= c[0] - 1 && lo[1] >= c[1] - 1 && lo[2] >= c[2] - 1 && hi[0] <= c[0] + 1 &&
                          hi[1] <= c[1] + 1 && hi[2] <= c[2] + 1;
        if (masks && inside && near) {
            uint32_t m = 0;
            for (int x = lo[0]; x <= hi[0]; x++)
                for (int y = lo[1]; y <= hi[1]; y++)
                    for (int z = lo[2]; z <= hi[2]; z++)
                        m |= 1u << ((z - c[2] + 1) * 9 + (y - c[1] + 1) * 3 + (x - c[0] + 1));
            m |= 1u << 13;
            uint32_t* wd = masks + idx;
            if ((__hip_atomic_load(wd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & m) != m)
                __hip_atomic_fetch_or(wd, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        if (inside) set_bit(bmA, idx);
        if (dilate) {
            for (int x = lo[0]; x <= hi[0]; x++)
                for (int y = lo[1]; y <= hi[1]; y++)
                    for (int z = lo[2]; z <= hi[2]; z++)
                        if (dom_index(d0, x, y, z, idx)) set_bit(bmB, idx);
        }
    }
}

// Direction masks -> level-0 bitmaps: bit 13 marks the brick itself in A (and
// in B), the other bits mark the neighbours in B.
__global__ __launch_bounds__(256) void k_expand_masks(BitDomain d0, const uint32_t* masks, uint32_t* bmA, uint32_t* bmB) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < d0.nbits; i += u64(gridDim.x) * blockDim.x) {
        uint32_t m = masks[i];
        if (!m) continue;
        int x, y, z;
        dom_coords(d0, i, x, y, z);
        if (m & (1u << 13)) set_bit(bmA, i);
        if (!bmB) continue;
        while (m) {
            int d = __ffs(m) - 1;
            m &= m - 1;
            u64 j;
            if (dom_index(d0, x + d % 3 - 1, y + (d / 3) % 3 - 1, z + d / 9 - 1, j)) set_bit(bmB, j);
        }
    }
}

hipError_t launch_mark_points(PointSrc pts, const float orig[3], float radius, int dilate, int range0, BitDomain d0,
                              uint32_t* bmA, uint32_t* bmB, uint32_t* masks, hipStream_t s) {
    int blocks = (pts.n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_mark_points, dim3(blocks), dim3(256), 0, s, pts, orig[0], orig[1], orig[2], radius, dilate,
                       range0, 1.0f / float(range0), d0, bmA, bmB, masks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !masks) return e;
    u64 b2 = (d0.nbits + 255) / 256;
    if (b2 > 16384) b2 = 16384;
    if (b2 < 1) b2 = 1;
    hipLaunchKernelGGL(k_expand_masks, dim3(unsigned(b2)), dim3(256), 0, s, d0, masks, bmA, dilate ? bmB : nullptr);
    return hipGetLastError();
}

// Mark level L+1 from the set bits of level L (parent = child >> shift).
__global__ __launch_bounds__(256) void k_mark_parents(BitDomain dc, const uint32_t* bmA_c, const uint32_t* bmB_c,
                                                      int shift, BitDomain dp, uint32_t* bmA_p, uint32_t* bmB_p) {
    for (u64 w = u64(blockIdx.x) * blockDim.x + threadIdx.x; w < dc.nwords; w += u64(gridDim.x) * blockDim.x) {
        for (int pass = 0; pass < 2; pass++) {
            uint32_t bits = pass == 0 ? bmA_c[w] : (bmB_c ? bmB_c[w] : 0u);
            uint32_t* out = pass == 0 ? bmA_p : bmB_p;
            u64 last = ~u64(0);
            while (bits) {
                int b = __ffs(bits) - 1;
                bits &= bits - 1;
                int x, y, z;
                dom_coords(dc, w * 32 + b, x, y, z);
                u64 idx;
                if (dom_index(dp, x >> shift, y >> shift, z >> shift, idx) && idx != last) {
                    set_bit(out, idx);
                    last = idx;
                }
            }
        }
    }
}

hipError_t launch_mark_parents(BitDomain dc, const uint32_t* bmA_c, const uint32_t* bmB_c, int shift, BitDomain dp,
                               uint32_t* bmA_p, uint32_t* bmB_p, hipStream_t s) {
    u64 blocks = (dc.nwords + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_mark_parents, dim3(unsigned(blocks)), dim3(256), 0, s, dc, bmA_c, bmB_c, shift, dp, bmA_p,
                       bmB_p);
    return hipGetLastError();
}

__device__ inline void write_node(Node* n, int lev, int x, int y, int z, int range, u64 clist) {
    Node v;
    v.mLev = uint8_t(lev);
    v.mFlags = 1;
    v.mPriority = 0;
    v.pad = 0;
    v.mPos[0] = x * range; v.mPos[1] = y * range; v.mPos[2] = z * range;
    v.mValue[0] = v.mValue[1] = v.mValue[2] = -1;
    v.mVRange[0] = v.mVRange[1] = v.mVRange[2] = 0.f;
    v.mParent = ID_UNDEFL;
    v.mChildList = clist;
    v.mMask = 0;
    *n = v;
}

// Node creation in sorted key order (src/gvdb_volume_gvdb.cpp:1308-1331 for the
// first pass, :965-988 for the appended dilation nodes).
__global__ __launch_bounds__(256) void k_emit_nodes(LevelBuild L, Node* nodes, int has_clist) {
    for (u64 w = u64(blockIdx.x) * blockDim.x + threadIdx.x; w < L.dom.nwords; w += u64(gridDim.x) * blockDim.x) {
        uint32_t a = L.bmA[w];
        uint32_t idx = L.rankA[w];
        uint32_t bits = a;
        while (bits) {
            int b = __ffs(bits) - 1;
            bits &= bits - 1;
            int x, y, z;
            dom_coords(L.dom, w * 32 + b, x, y, z);
            write_node(nodes + idx, L.lev, x, y, z, L.range, has_clist ? Elem(1, L.lev, idx) : ID_UNDEFL);
            idx++;
        }
        if (L.bmB != L.bmA && L.bmB) {
            uint32_t e = L.bmB[w] & ~a;
            uint32_t j = L.countA + L.rankE[w];
            while (e) {
                int b = __ffs(e) - 1;
                e &= e - 1;
                int x, y, z;
                dom_coords(L.dom, w * 32 + b, x, y, z);
                write_node(nodes + j, L.lev, x, y, z, L.range, has_clist ? Elem(1, L.lev, j) : ID_UNDEFL);
                j++;
            }
        }
    }
}

hipError_t launch_emit_nodes(const LevelBuild& L, Node* nodes, int has_clist, hipStream_t s) {
    u64 blocks = (L.dom.nwords + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_emit_nodes, dim3(unsigned(blocks)), dim3(256), 0, s, L, nodes, has_clist);
    return hipGetLastError();
}

__device__ inline uint32_t rank_of(const LevelBuild& P, u64 idx) {
    u64 w = idx >> 5;
    uint32_t m = (1u << (idx & 31)) - 1u;
    uint32_t a = P.bmA[w];
    if (a & (1u << (idx & 31))) return P.rankA[w] + __popc(a & m);
    if (!P.rankE || P.bmB == P.bmA) return P.countA;   // not reachable for nodes of a valid build
    uint32_t e = P.bmB[w] & ~a;
    return P.countA + P.rankE[w] + __popc(e & m);
}

// gvdbLinkBricks (kernels/cuda_gvdb_nodes.cuh:470-505): parent pointer and
// the child-list slot (z*res + y)*res + x of the node inside its parent.
__global__ __launch_bounds__(256) void k_link_level(LevelBuild C, LevelBuild P, int parent_is_root, int logdim_p,
                                                    Node* cnodes, u64* pclists, u64 clist_len) {
    uint32_t total = C.countA + C.countE;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        Node* n = cnodes + i;
        int c[3] = {n->mPos[0] / C.range, n->mPos[1] / C.range, n->mPos[2] / C.range};
        int pc[3] = {c[0] >> logdim_p, c[1] >> logdim_p, c[2] >> logdim_p};
        uint32_t pidx = 0;
        if (!parent_is_root) {
            u64 idx;
            if (!dom_index(P.dom, pc[0], pc[1], pc[2], idx)) continue;
            pidx = rank_of(P, idx);
        }
        int res = 1 << logdim_p;
        int l[3] = {c[0] - (pc[0] << logdim_p), c[1] - (pc[1] << logdim_p), c[2] - (pc[2] << logdim_p)};
        int bit = (l[2] * res + l[1]) * res + l[0];
        n->mParent = Elem(0, C.lev + 1, pidx);
        pclists[u64(pidx) * clist_len + bit] = Elem(0, C.lev, i);
    }
}

hipError_t launch_link_level(const LevelBuild& child, const LevelBuild* parent, int parent_is_root, int logdim_parent,
                             Node* cnodes, u64* parent_clists, u64 clist_len, hipStream_t s) {
    uint32_t total = child.countA + child.countE;
    if (total == 0) return hipSuccess;
    unsigned blocks = (total + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    LevelBuild P = parent ? *parent : child;
    hipLaunchKernelGGL(k_link_level, dim3(blocks), dim3(256), 0, s, child, P, parent_is_root, logdim_parent, cnodes,
                       parent_clists, clist_len);
    return hipGetLastError();
}

__global__ void k_write_root(Node* root, int x, int y, int z, int lev) {
    Node v;
    v.mLev = uint8_t(lev);
    v.mFlags = 1;
    v.mPriority = 0;
    v.pad = 0;
    v.mPos[0] = x; v.mPos[1] = y; v.mPos[2] = z;
    v.mValue[0] = v.mValue[1] = v.mValue[2] = -1;
    v.mVRange[0] = v.mVRange[1] = v.mVRange[2] = 0.f;
    v.mParent = ID_UNDEFL;
    v.mChildList = Elem(1, lev, 0);
    v.mMask = 0;
    *root = v;
}

hipError_t launch_write_root(Node* root, const int pos[3], int lev, hipStream_t s) {
    hipLaunchKernelGGL(k_write_root, dim3(1), dim3(1), 0, s, root, pos[0], pos[1], pos[2], lev);
    return hipGetLastError();
}

}  // namespace gvdb
