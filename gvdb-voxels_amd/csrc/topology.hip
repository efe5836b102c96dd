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
#include "tree_device.h"

namespace gvdb {

__device__ inline uint32_t float_key(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ inline const float* point_at(const PointSrc& p, int i) {
    return reinterpret_cast<const float*>(p.base + u64(i) * p.stride + p.off);
}

static uint32_t grid_for(u64 items, uint32_t cap) {
    u64 b = (items + 255) / 256;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return uint32_t(b);
}

// GetBoundingBox's split + Thrust min/max (src/gvdb_volume_gvdb.cpp:5404-5467).
__device__ inline void k_bbox_body(const BboxArgs& a, uint32_t bid, uint32_t nblk) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    const PointSrc& P = a.pts;
    if (P.stride == 12 && P.off == 0 && P.n < (1 << 30) && (reinterpret_cast<uintptr_t>(P.base) & 15) == 0) {
        // packed float3: four points = three float4 per thread (axes fixed:
        // x y z x | y z x y | z x y z), then the last n % 4 points
        const float4* v = reinterpret_cast<const float4*>(P.base);
        const uint32_t ng = uint32_t(P.n) / 4u;
        const float ox = a.o[0], oy = a.o[1], oz = a.o[2];
        auto ux = [&](float f) { f += ox; mn[0] = fminf(mn[0], f); mx[0] = fmaxf(mx[0], f); };
        auto uy = [&](float f) { f += oy; mn[1] = fminf(mn[1], f); mx[1] = fmaxf(mx[1], f); };
        auto uz = [&](float f) { f += oz; mn[2] = fminf(mn[2], f); mx[2] = fmaxf(mx[2], f); };
        for (uint32_t g = bid * 256 + threadIdx.x; g < ng; g += nblk * 256) {
            const float4 q0 = v[3 * u64(g)], q1 = v[3 * u64(g) + 1], q2 = v[3 * u64(g) + 2];
            ux(q0.x); uy(q0.y); uz(q0.z); ux(q0.w);
            uy(q1.x); uz(q1.y); ux(q1.z); uy(q1.w);
            uz(q2.x); ux(q2.y); uy(q2.z); uz(q2.w);
        }
        if (bid == 0 && threadIdx.x < uint32_t(P.n) - 4u * ng) {
            const float* p = point_at(P, int(4u * ng + threadIdx.x));
            ux(p[0]); uy(p[1]); uz(p[2]);
        }
    } else {
        for (int i = bid * 256 + threadIdx.x; i < P.n; i += nblk * 256) {
            const float* p = point_at(P, i);
            float w[3] = {p[0] + a.o[0], p[1] + a.o[1], p[2] + a.o[2]};
#pragma unroll
            for (int k = 0; k < 3; k++) { mn[k] = fminf(mn[k], w[k]); mx[k] = fmaxf(mx[k], w[k]); }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            mn[k] = fminf(mn[k], __shfl_xor(mn[k], d, 64));
            mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], d, 64));
        }
    }
    __shared__ float smn[4][3], smx[4][3];
    int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; k++) { smn[wave][k] = mn[k]; smx[wave][k] = mx[k]; }
    __syncthreads();
    // each block stores its six keys; k_bbox_reduce combines them (a kernel
    // boundary orders the stores: no same-address atomics, no fences)
    if (threadIdx.x < 3) {
        int k = threadIdx.x;
        float m0 = smn[0][k], m1 = smx[0][k];
        for (int w = 1; w < 4; w++) { m0 = fminf(m0, smn[w][k]); m1 = fmaxf(m1, smx[w][k]); }
        a.part[bid * 6 + k] = float_key(m0);
        a.part[bid * 6 + 3 + k] = float_key(m1);
    }
}
GVDB_MULTI(k_bbox, BboxArgs, 256)

// One block per grid: the minimum and maximum over the blocks' keys.
__device__ inline void k_bbox_reduce_body(const BboxArgs& a, uint32_t, uint32_t) {
    uint32_t k6[6] = {~0u, ~0u, ~0u, 0u, 0u, 0u};
    for (uint32_t j = threadIdx.x; j < a.nparts; j += 256) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            k6[k] = min(k6[k], a.part[j * 6 + k]);
            k6[3 + k] = max(k6[3 + k], a.part[j * 6 + 3 + k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            k6[k] = min(k6[k], uint32_t(__shfl_xor(int(k6[k]), d, 64)));
            k6[3 + k] = max(k6[3 + k], uint32_t(__shfl_xor(int(k6[3 + k]), d, 64)));
        }
    }
    __shared__ uint32_t sk[4][6];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; k++) sk[wave][k] = k6[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        uint32_t m = sk[0][k];
        for (int w = 1; w < 4; w++) m = k < 3 ? min(m, sk[w][k]) : max(m, sk[w][k]);
        a.keys6[k] = m;
    }
}
GVDB_MULTI(k_bbox_reduce, BboxArgs, 256)


hipError_t launch_bbox_multi(const BboxArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<BboxArgs> r(a, a + G);
    std::vector<uint32_t> nb(G), one(G, 1u);
    for (int g = 0; g < G; g++) {
        nb[g] = grid_for(u64(a[g].pts.n) / 4, BBOX_BLOCKS);
        r[g].nparts = nb[g];
    }
    MultiLaunch L;
    hipError_t e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    hipLaunchKernelGGL(k_bbox, dim3(L.blocks), dim3(256), 0, s, a[0], L.map);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = multi_setup(ar, r.data(), one.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    hipLaunchKernelGGL(k_bbox_reduce, dim3(L.blocks), dim3(256), 0, s, r[0], L.map);
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

// Bits are only ever set while a bitmap is built, so the plain (L2-cached)
// read that filters the OR can only miss bits set meanwhile by other XCDs;
// then the OR is done anyway.
__device__ inline void set_bit(uint32_t* bm, u64 idx) {
    uint32_t m = 1u << (idx & 31);
    uint32_t* w = bm + (idx >> 5);
    if (!(*w & m)) __hip_atomic_fetch_or(w, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// gvdbCalcBrickId (kernels/cuda_gvdb_particles.cuh:352-382) for level 0 and
// gvdbCalcExtraBrickId (:467-512) for the dilation.  Each point contributes
// its brick and, for the dilation, the bricks covering [int(p)-r, int(p)+2r]
// on every axis.  With `bytes` (domains up to 2^27 bricks) each mark is a
// byte store of 1 after a plain read (no atomics: concurrent stores write the
// same value), packed into the bitmaps by k_pack_bytes.  Otherwise:  When those lie within one brick of the point's own (r up to
// one brick), the point ORs a 27-bit direction mask into its own brick's word
// (`masks`, bit 13 = the brick itself), which k_expand_masks turns into bitmap
// bits; this replaces up to 8 bitmap atomics per point by one rarely taken
// atomic.  Points with wider reach (or no mask array) set the bits directly.
__device__ inline void k_mark_points_body(const MarkArgs& a, uint32_t bid, uint32_t nblk) {
    const BitDomain d0 = a.d0;
    const int range0 = a.range0;
    for (int i = bid * 256 + threadIdx.x; i < a.pts.n; i += nblk * 256) {
        const float* p = point_at(a.pts, i);
        float w[3] = {p[0] + a.o[0], p[1] + a.o[1], p[2] + a.o[2]};
        if (!isfinite(w[0]) || !isfinite(w[1]) || !isfinite(w[2])) continue;
        if (w[0] < 0 || w[1] < 0 || w[2] < 0) continue;
        int c[3];
#pragma unroll
        for (int k = 0; k < 3; k++) c[k] = int(floorf(w[k] * a.inv_range0));
        u64 idx;
        const bool inside = dom_index(d0, c[0], c[1], c[2], idx);
        int lo[3] = {c[0], c[1], c[2]}, hi[3] = {c[0], c[1], c[2]};
        if (a.dilate) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                int iw = int(w[k]);
                int bmin = int(float(iw) - a.radius);
                int bmax = int(float(iw) + 2.0f * a.radius);
                lo[k] = bmin / range0;   // C truncation, as the reference's int division
                hi[k] = bmax / range0;
            }
        }
        if (a.bytes) {
            // one plain byte store per marked brick: concurrent stores of the
            // same value need no atomics; k_pack_bytes turns bytes into bits
            // a plain read skips the store where the byte is already set
            if (inside && !a.bytes[idx]) a.bytes[idx] = 1;
            if (a.dilate) {
                uint8_t* bb = a.bytes + d0.nwords * 32;
                for (int x = lo[0]; x <= hi[0]; x++)
                    for (int y = lo[1]; y <= hi[1]; y++)
                        for (int z = lo[2]; z <= hi[2]; z++)
                            if (dom_index(d0, x, y, z, idx) && !bb[idx]) bb[idx] = 1;
            }
            continue;
        }
        const bool near = lo[0] >= c[0] - 1 && lo[1] >= c[1] - 1 && lo[2] >= c[2] - 1 && hi[0] <= c[0] + 1 &&
                          hi[1] <= c[1] + 1 && hi[2] <= c[2] + 1;
        if (a.masks && inside && near) {
            uint32_t m = 0;
            for (int x = lo[0]; x <= hi[0]; x++)
                for (int y = lo[1]; y <= hi[1]; y++)
                    for (int z = lo[2]; z <= hi[2]; z++)
                        m |= 1u << ((z - c[2] + 1) * 9 + (y - c[1] + 1) * 3 + (x - c[0] + 1));
            m |= 1u << 13;
            uint32_t* wd = a.masks + idx;
            // Bits are only ever set here, so a plain (L2-cached) read can
            // only miss bits other XCDs set since: then the OR is done anyway.
            // Skipping the OR needs every bit of m already present.
            if ((*wd & m) != m) __hip_atomic_fetch_or(wd, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        if (inside) set_bit(a.bmA, idx);
        if (a.dilate) {
            for (int x = lo[0]; x <= hi[0]; x++)
                for (int y = lo[1]; y <= hi[1]; y++)
                    for (int z = lo[2]; z <= hi[2]; z++)
                        if (dom_index(d0, x, y, z, idx)) set_bit(a.bmB, idx);
        }
    }
}
GVDB_MULTI(k_mark_points, MarkArgs, 256)

// Direction masks -> level-0 bitmaps: bit 13 marks the brick itself in A (and
// in B), the other bits mark the neighbours in B.
__device__ inline void k_expand_masks_body(const MarkArgs& a, uint32_t bid, uint32_t nblk) {
    const BitDomain d0 = a.d0;
    uint32_t* bmB = a.dilate ? a.bmB : nullptr;
    for (u64 i = u64(bid) * 256 + threadIdx.x; i < d0.nbits; i += u64(nblk) * 256) {
        uint32_t m = a.masks[i];
        if (!m) continue;
        int x, y, z;
        dom_coords(d0, i, x, y, z);
        if (m & (1u << 13)) set_bit(a.bmA, i);
        if (!bmB) continue;
        while (m) {
            int d = __ffs(m) - 1;
            m &= m - 1;
            u64 j;
            if (dom_index(d0, x + d % 3 - 1, y + (d / 3) % 3 - 1, z + d / 9 - 1, j)) set_bit(bmB, j);
        }
    }
}
GVDB_MULTI(k_expand_masks, MarkArgs, 256)

// Byte marks -> level-0 bitmaps: 32 bytes (each 0 or 1) per 32-bit word.
__device__ inline uint32_t pack32(const uint8_t* b) {
    const uint4* q = reinterpret_cast<const uint4*>(b);
    uint32_t m = 0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint4 v = q[h];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t x = w[k];
            m |= ((x & 1u) | ((x >> 7) & 2u) | ((x >> 14) & 4u) | ((x >> 21) & 8u)) << (16 * h + 4 * k);
        }
    }
    return m;
}

__device__ inline void k_pack_bytes_body(const MarkArgs& a, uint32_t bid, uint32_t nblk) {
    const BitDomain d0 = a.d0;
    for (u64 w = u64(bid) * 256 + threadIdx.x; w < d0.nwords; w += u64(nblk) * 256) {
        a.bmA[w] = pack32(a.bytes + w * 32);
        if (a.dilate) a.bmB[w] = pack32(a.bytes + d0.nwords * 32 + w * 32);
    }
}
GVDB_MULTI(k_pack_bytes, MarkArgs, 256)

hipError_t launch_mark_points_multi(const MarkArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> nb(G), nb2(G), nb3(G);
    bool any_mask = false, any_bytes = false;
    for (int g = 0; g < G; g++) {
        nb[g] = a[g].pts.n > 0 ? grid_for(u64(a[g].pts.n), 16384) : 0;
        nb2[g] = a[g].masks && !a[g].bytes ? grid_for(a[g].d0.nbits, 16384) : 0;
        nb3[g] = a[g].bytes ? grid_for(a[g].d0.nwords, 16384) : 0;
        any_mask |= a[g].masks != nullptr && !a[g].bytes;
        any_bytes |= a[g].bytes != nullptr;
    }
    MultiLaunch L;
    hipError_t e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess) return e;
    if (L.blocks) {
        hipLaunchKernelGGL(k_mark_points, dim3(L.blocks), dim3(256), 0, s, a[0], L.map);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (any_bytes) {
        MultiLaunch L3;
        e = multi_setup(ar, a, nb3.data(), G, s, L3);
        if (e != hipSuccess) return e;
        if (L3.blocks) {
            hipLaunchKernelGGL(k_pack_bytes, dim3(L3.blocks), dim3(256), 0, s, a[0], L3.map);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    if (!any_mask) return hipSuccess;
    MultiLaunch L2;
    e = multi_setup(ar, a, nb2.data(), G, s, L2);
    if (e != hipSuccess || L2.blocks == 0) return e;
    hipLaunchKernelGGL(k_expand_masks, dim3(L2.blocks), dim3(256), 0, s, a[0], L2.map);
    return hipGetLastError();
}

// Node emission walks domains up to this many bits one thread per bit (dense
// words would serialise a thread-per-word loop); larger, sparser ones per word.
static const u64 PER_BIT_MAX = BRICK_TAB_MAX;   // one thread per bit (and the brick table) up to here

// Mark level L+1 from the set bits of level L (parent = child >> shift):
// one wave per parent bit reads its children's rows (lane = one (x, y) row of
// 2^shift z bits, consecutive in x-major order) and one lane sets the bit,
// so no two threads race on a parent word's atomics more than once per
// parent.
__device__ inline bool any_bits(const uint32_t* bm, u64 i0, u64 i1) {
    // any bit set in [i0, i1] (inclusive)
    for (u64 w = i0 >> 5; w <= (i1 >> 5); w++) {
        uint32_t m = ~0u;
        if (w == (i0 >> 5)) m &= ~0u << (i0 & 31);
        if (w == (i1 >> 5)) m &= ~0u >> (31 - (i1 & 31));
        if (bm[w] & m) return true;
    }
    return false;
}

__device__ inline void k_mark_parents_body(const ParentArgs& a, uint32_t bid, uint32_t nblk) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int s = a.shift, side = 1 << s;
    const BitDomain& c = a.dc;
    for (u64 p = u64(bid) * 4 + wave; p < a.dp.nbits; p += u64(nblk) * 4) {
        int px, py, pz;
        dom_coords(a.dp, p, px, py, pz);
        const int z0 = max(pz << s, c.lo[2]), z1 = min(((pz + 1) << s) - 1, c.lo[2] + c.dim[2] - 1);
        bool hasA = false, hasB = false;
        if (z0 <= z1) {
            for (int r = lane; r < side * side; r += 64) {
                const int cx = (px << s) + (r & (side - 1)), cy = (py << s) + (r >> s);
                if (cx < c.lo[0] || cy < c.lo[1] || cx >= c.lo[0] + c.dim[0] || cy >= c.lo[1] + c.dim[1]) continue;
                const u64 i0 = (u64(cx - c.lo[0]) * c.dim[1] + u64(cy - c.lo[1])) * c.dim[2] + u64(z0 - c.lo[2]);
                const u64 i1 = i0 + u64(z1 - z0);
                hasA = hasA || any_bits(a.Ac, i0, i1);
                if (a.Bc) hasB = hasB || any_bits(a.Bc, i0, i1);
            }
        }
        const u64 ba = __ballot(hasA), bb = __ballot(hasB);
        if (lane == 0) {
            if (ba) set_bit(a.Ap, p);
            if (bb && a.Bp) set_bit(a.Bp, p);
        }
    }
}
GVDB_MULTI(k_mark_parents, ParentArgs, 256)

hipError_t launch_mark_parents_multi(const ParentArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> nb(G);
    for (int g = 0; g < G; g++) nb[g] = uint32_t(std::min<u64>((a[g].dp.nbits + 3) / 4, 65536));
    MultiLaunch L;
    hipError_t e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    hipLaunchKernelGGL(k_mark_parents, dim3(L.blocks), dim3(256), 0, s, a[0], L.map);
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
__device__ inline void k_emit_nodes_body(const EmitArgs& a, uint32_t bid, uint32_t nblk) {
    const LevelBuild& L = a.L;
    if (L.dom.nbits <= PER_BIT_MAX) {
        const bool dil = L.bmB != L.bmA && L.bmB;
        for (u64 i = u64(bid) * 256 + threadIdx.x; i < L.dom.nbits; i += u64(nblk) * 256) {
            const u64 w = i >> 5;
            const uint32_t bit = 1u << (i & 31), below = bit - 1u;
            const uint32_t av = L.bmA[w];
            uint32_t idx = ~0u;
            if (av & bit) {
                idx = L.rankA[w] + uint32_t(__popc(av & below));
            } else {
                const uint32_t e = dil ? (L.bmB[w] & ~av) : 0u;
                if (e & bit) idx = L.countA + L.rankE[w] + uint32_t(__popc(e & below));
            }
            if (a.btab) a.btab[i] = idx;   // every bit of the box (mask word 0), so no clear is needed
            if (idx == ~0u) continue;
            int x, y, z;
            dom_coords(L.dom, i, x, y, z);
            write_node(a.nodes + idx, L.lev, x, y, z, L.range, a.has_clist ? Elem(1, L.lev, idx) : ID_UNDEFL);
        }
        return;
    }
    for (u64 w = u64(bid) * 256 + threadIdx.x; w < L.dom.nwords; w += u64(nblk) * 256) {
        uint32_t av = L.bmA[w];
        uint32_t idx = L.rankA[w];
        uint32_t bits = av;
        while (bits) {
            int b = __ffs(bits) - 1;
            bits &= bits - 1;
            int x, y, z;
            dom_coords(L.dom, w * 32 + b, x, y, z);
            write_node(a.nodes + idx, L.lev, x, y, z, L.range, a.has_clist ? Elem(1, L.lev, idx) : ID_UNDEFL);
            idx++;
        }
        if (L.bmB != L.bmA && L.bmB) {
            uint32_t e = L.bmB[w] & ~av;
            uint32_t j = L.countA + L.rankE[w];
            while (e) {
                int b = __ffs(e) - 1;
                e &= e - 1;
                int x, y, z;
                dom_coords(L.dom, w * 32 + b, x, y, z);
                write_node(a.nodes + j, L.lev, x, y, z, L.range, a.has_clist ? Elem(1, L.lev, j) : ID_UNDEFL);
                j++;
            }
        }
    }
}
GVDB_MULTI(k_emit_nodes, EmitArgs, 256)

hipError_t launch_emit_nodes_multi(const EmitArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> nb(G);
    for (int g = 0; g < G; g++)
        nb[g] = grid_for(a[g].L.dom.nbits <= PER_BIT_MAX ? a[g].L.dom.nbits : a[g].L.dom.nwords, 16384);
    MultiLaunch L;
    hipError_t e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    hipLaunchKernelGGL(k_emit_nodes, dim3(L.blocks), dim3(256), 0, s, a[0], L.map);
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
__device__ inline void k_link_level_body(const LinkArgs& a, uint32_t bid, uint32_t nblk) {
    const LevelBuild& C = a.C;
    const uint32_t total = C.countA + C.countE;
    const int lp = a.logdim_p;
    for (uint32_t i = bid * 256 + threadIdx.x; i < total; i += nblk * 256) {
        Node* n = a.cnodes + i;
        int c[3] = {n->mPos[0] / C.range, n->mPos[1] / C.range, n->mPos[2] / C.range};
        int pc[3] = {c[0] >> lp, c[1] >> lp, c[2] >> lp};
        uint32_t pidx = 0;
        if (!a.parent_is_root) {
            u64 idx;
            if (!dom_index(a.P.dom, pc[0], pc[1], pc[2], idx)) continue;
            pidx = rank_of(a.P, idx);
        }
        int res = 1 << lp;
        int l[3] = {c[0] - (pc[0] << lp), c[1] - (pc[1] << lp), c[2] - (pc[2] << lp)};
        int bit = (l[2] * res + l[1]) * res + l[0];
        n->mParent = Elem(0, C.lev + 1, pidx);
        a.pclists[u64(pidx) * a.clist_len + bit] = Elem(0, C.lev, i);
    }
}
GVDB_MULTI(k_link_level, LinkArgs, 256)

hipError_t launch_link_level_multi(const LinkArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> nb(G);
    for (int g = 0; g < G; g++) {
        u64 t = u64(a[g].C.countA) + a[g].C.countE;
        nb[g] = t ? grid_for(t, 16384) : 0;
    }
    MultiLaunch L;
    hipError_t e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    hipLaunchKernelGGL(k_link_level, dim3(L.blocks), dim3(256), 0, s, a[0], L.map);
    return hipGetLastError();
}

// Root node (SetupNode at the top level, src/gvdb_volume_gvdb.cpp:1289-1300).
__device__ inline void k_write_root_body(const RootArgs& a, uint32_t, uint32_t) {
    if (threadIdx.x != 0) return;
    Node v;
    v.mLev = uint8_t(a.lev);
    v.mFlags = 1;
    v.mPriority = 0;
    v.pad = 0;
    v.mPos[0] = a.pos[0]; v.mPos[1] = a.pos[1]; v.mPos[2] = a.pos[2];
    v.mValue[0] = v.mValue[1] = v.mValue[2] = -1;
    v.mVRange[0] = v.mVRange[1] = v.mVRange[2] = 0.f;
    v.mParent = ID_UNDEFL;
    v.mChildList = Elem(1, a.lev, 0);
    v.mMask = 0;
    *a.root = v;
}
GVDB_MULTI(k_write_root, RootArgs, 64)

hipError_t launch_write_root_multi(const RootArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> nb(G, 1u);
    MultiLaunch L;
    hipError_t e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    hipLaunchKernelGGL(k_write_root, dim3(L.blocks), dim3(64), 0, s, a[0], L.map);
    return hipGetLastError();
}


// ------------------------------------------------- incremental topology ----
// ActivateIncreBricksGPU (src/gvdb_volume_gvdb.cpp:1020-1156).  The reference
// appends the keys of missing nodes with an atomic counter and sorts them;
// here they are bits of one x-major bitmap per level (bit order = key order),
// so a scan gives each new node its place.

// gvdbCalcIncreExtraBrickId (kernels/cuda_gvdb_particles.cuh:420-465): the
// box is [int(int(w) - r), int(int(w) + r)] per axis, node corners from
// (C-truncated) box / range * range.  Corners below zero (radius 8 or more)
// are skipped: the reference would cast them into its u16 keys.
__global__ __launch_bounds__(256) void k_incre_mark(IncreMarkArgs a) {
    const Tree& t = a.t;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < a.pts.n; i += gridDim.x * 256) {
        const float* p = point_at(a.pts, i);
        const float w[3] = {p[0] + a.o[0], p[1] + a.o[1], p[2] + a.o[2]};
        if (!isfinite(w[0]) || !isfinite(w[1]) || !isfinite(w[2])) continue;
        if (w[0] < 0 || w[1] < 0 || w[2] < 0) continue;
        int bmin[3], bmax[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int iw = int(w[k]);
            bmin[k] = int(float(iw) - a.radius);
            bmax[k] = int(float(iw) + a.radius);
        }
        for (int lev = 0; lev < a.D; lev++) {
            const int rr = t.range[lev];
            const int x0 = bmin[0] / rr * rr, x1 = bmax[0] / rr * rr;
            const int y0 = bmin[1] / rr * rr, y1 = bmax[1] / rr * rr;
            const int z0 = bmin[2] / rr * rr, z1 = bmax[2] / rr * rr;
            for (int x = x0; x <= x1; x += rr)
                for (int y = y0; y <= y1; y += rr)
                    for (int z = z0; z <= z1; z += rr) {
                        if (x < 0 || y < 0 || z < 0) continue;
                        const int id = pos_parent(t, x, y, z, lev);
                        if (id < 0) {
                            u64 idx;
                            if (dom_index(a.dom[lev], x / rr, y / rr, z / rr, idx)) set_bit(a.bm[lev], idx);
                        } else if (lev == 0) {
                            a.marker[id] = 1u;
                        }
                    }
        }
    }
}

hipError_t launch_incre_mark(const IncreMarkArgs& a, hipStream_t s) {
    if (a.pts.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_incre_mark, dim3(grid_for(u64(a.pts.n), 16384)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// "Update markers": the leaves present before the call take their markers.
__global__ __launch_bounds__(256) void k_apply_markers(Node* leaves, const uint32_t* marker, u64 n) {
    const u64 i = u64(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) leaves[i].mFlags = uint8_t(marker[i]);
}

hipError_t launch_apply_markers(Node* leaves, const uint32_t* marker, u64 n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_apply_markers, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, leaves, marker, n);
    return hipGetLastError();
}

// SetupNode + child-list allocation of the appended nodes
// (src/gvdb_volume_gvdb.cpp:1087-1110), in key order.
__global__ __launch_bounds__(256) void k_incre_emit(IncreEmitArgs a) {
    for (u64 w = u64(blockIdx.x) * 256 + threadIdx.x; w < a.dom.nwords; w += u64(gridDim.x) * 256) {
        uint32_t bits = a.bm[w];
        u64 r = a.rank[w];
        while (bits) {
            const int b = __ffs(bits) - 1;
            bits &= bits - 1;
            int x, y, z;
            dom_coords(a.dom, w * 32 + b, x, y, z);
            write_node(a.nodes + a.base + r, a.lev, x, y, z, a.range,
                       a.has_clist ? Elem(1, a.lev, a.cbase + r) : ID_UNDEFL);
            r++;
        }
    }
}

hipError_t launch_incre_emit(const IncreEmitArgs& a, hipStream_t s) {
    if (a.dom.nwords == 0) return hipSuccess;
    hipLaunchKernelGGL(k_incre_emit, dim3(grid_for(a.dom.nwords, 16384)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Cell of `pos` inside the node `pn` of level plev: ((pos - corner) * res) /
// range per axis, as the reference computes it.
__device__ inline bool cell_in(const Tree& t, int plev, const Node* pn, const int pos[3], int& bit) {
    const int res = t.res[plev];
    const int sh = __builtin_ctz(unsigned(t.range[plev])) - t.logdim[plev];
    int c[3];
#pragma unroll
    for (int k = 0; k < 3; k++) c[k] = child_cell((long long)pos[k] - pn->mPos[k], sh);
    if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= res || c[1] >= res || c[2] >= res) return false;
    bit = (c[2] * res + c[1]) * res + c[0];
    return true;
}

// gvdbLinkBricks (kernels/cuda_gvdb_nodes.cuh:470-505) over every node of
// one level: flagged nodes are put in the slot of the node one level up
// that the walk from the top node reaches.  The walk reads child lists two
// or more levels up, the writes go one level up, so one launch per level.
__global__ __launch_bounds__(256) void k_link_walk(Tree t, int lev) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= t.nodecnt[lev]) return;
    Node* n = t.nodes[lev] + i;
    if (!n->mFlags) return;
    const int pid = get_parent(t, lev + 1, n->mPos);
    if (pid < 0) return;
    const Node* pn = t.nodes[lev + 1] + pid;
    int bit;
    if (!cell_in(t, lev + 1, pn, n->mPos, bit)) return;
    n->mParent = Elem(0, lev + 1, u64(pid));
    const u64 cndx = pn->mChildList >> 16;
    if (cndx >= u64(t.nodecnt[lev + 1])) return;
    const u64 r = u64(t.res[lev + 1]);
    t.clists[lev + 1][cndx * r * r * r + u64(bit)] = Elem(0, lev, u64(i));
}

hipError_t launch_link_walk(const Tree& t, int lev, hipStream_t s) {
    const int n = t.nodecnt[lev];
    if (n <= 0 || lev + 1 >= t.levels || !t.clists[lev + 1]) return hipSuccess;
    hipLaunchKernelGGL(k_link_walk, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, t, lev);
    return hipGetLastError();
}

// Empties the parent's slot of node n (level lev) and marks n parentless.
__device__ inline void delink_from_parent(const Tree& t, int lev, Node* n) {
    const int pid = get_parent(t, lev + 1, n->mPos);
    if (pid < 0) return;
    const Node* pn = t.nodes[lev + 1] + pid;
    int bit;
    if (!cell_in(t, lev + 1, pn, n->mPos, bit)) return;
    n->mParent = ID_UNDEF64;
    const u64 r = u64(t.res[lev + 1]);
    t.clists[lev + 1][(pn->mChildList >> 16) * r * r * r + u64(bit)] = ID_UNDEF64;
}

// gvdbDelinkLeafBricks (kernels/cuda_gvdb_nodes.cuh:401-429)
__global__ __launch_bounds__(256) void k_delink_leaves(Tree t) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= t.nodecnt[0]) return;
    Node* n = t.nodes[0] + i;
    if (n->mFlags || n->mParent == ID_UNDEF64) return;
    delink_from_parent(t, 0, n);
}

hipError_t launch_delink_leaves(const Tree& t, hipStream_t s) {
    const int n = t.nodecnt[0];
    if (n <= 0 || t.levels < 2 || !t.clists[1]) return hipSuccess;
    hipLaunchKernelGGL(k_delink_leaves, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, t);
    return hipGetLastError();
}

// gvdbDelinkBricks (kernels/cuda_gvdb_nodes.cuh:431-467): a parentless node
// is unflagged; a node with a child is flagged; a childless one is unflagged
// and leaves its parent.  Its own list is read at its pool index.
__global__ __launch_bounds__(256) void k_delink_level(Tree t, int lev) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= t.nodecnt[lev]) return;
    Node* n = t.nodes[lev] + i;
    if (n->mParent == ID_UNDEF64) { n->mFlags = 0; return; }
    const u64 r = u64(t.res[lev]);
    const u64 len = r * r * r;
    const u64* cl = t.clists[lev] + u64(i) * len;
    for (u64 c = 0; c < len; c++)
        if (cl[c] != ID_UNDEF64) { n->mFlags = 1; return; }
    n->mFlags = 0;
    delink_from_parent(t, lev, n);
}

hipError_t launch_delink_level(const Tree& t, int lev, hipStream_t s) {
    const int n = t.nodecnt[lev];
    if (n <= 0 || lev + 1 >= t.levels || !t.clists[lev] || !t.clists[lev + 1]) return hipSuccess;
    hipLaunchKernelGGL(k_delink_level, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, t, lev);
    return hipGetLastError();
}

// UpdateUsed (src/gvdb_volume_gvdb.cpp:1145-1149): unflagged leaves.
__global__ __launch_bounds__(256) void k_count_unflagged(Tree t, uint32_t* out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool off = i < t.nodecnt[0] && !t.nodes[0][i].mFlags;
    const u64 b = __ballot(off);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(out, uint32_t(__popcll(b)));
}

hipError_t launch_count_unflagged(const Tree& t, uint32_t* out, hipStream_t s) {
    const int n = t.nodecnt[0];
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_count_unflagged, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, t, out);
    return hipGetLastError();
}

}  // namespace gvdb
