// points.hip -- atlas slot assignment, subcell binning and the density gather.
#include "kernels.h"
#include <type_traits>
#include "tree_device.h"
#include <hip/hip_fp16.h>

namespace gvdb {

__device__ inline const float* point_at(const PointSrc& p, int i) {
    return reinterpret_cast<const float*>(p.base + u64(i) * p.stride + p.off);
}

// ------------------------------------------------------------ UpdateAtlas
// Slot per flagged leaf in pool order, mValue = slot corner + apron, and the
// atlas map entry (src/gvdb_volume_gvdb.cpp:2630-2704, getAtlasPos
// src/gvdb_allocator.cpp:705-715).  The map is cleared to -1 beforehand.
__device__ inline void k_update_atlas_body(const AtlasArgs& a, uint32_t bid, uint32_t) {
    const Tree& t = a.t;
    const uint32_t* slot_scan = a.slot_scan;
    int i = bid * 256 + threadIdx.x;
    if (i >= a.n) return;
    Node* n = t.nodes[0] + i;
    if (!n->mFlags) return;
    u64 id = slot_scan[i];
    u64 a2 = u64(t.atlas_cnt[0]) * t.atlas_cnt[1];
    int pz = int(id / a2);
    id -= u64(pz) * a2;
    int py = int(id / t.atlas_cnt[0]);
    id -= u64(py) * t.atlas_cnt[0];
    int px = int(id);
    int br = t.brick_res;
    n->mValue[0] = px * br + t.apron;
    n->mValue[1] = py * br + t.apron;
    n->mValue[2] = pz * br + t.apron;
    u64 slot = (u64(pz) * t.atlas_cnt[1] + py) * t.atlas_cnt[0] + px;
    AtlasNode an;
    an.mPos[0] = n->mPos[0]; an.mPos[1] = n->mPos[1]; an.mPos[2] = n->mPos[2];
    an.mLeafNode = i;
    t.amap[slot] = an;
}

GVDB_MULTI(k_update_atlas, AtlasArgs, 256)

template <typename A, void (*K)(A, MultiMap)>
static hipError_t launch_per_item(const A* a, const uint32_t* items, int G, Arena* ar, hipStream_t s, int threads) {
    std::vector<uint32_t> nb(G);
    for (int g = 0; g < G; g++) nb[g] = (items[g] + threads - 1) / threads;
    MultiLaunch L;
    hipError_t e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    hipLaunchKernelGGL(K, dim3(L.blocks), dim3(threads), 0, s, a[0], L.map);
    return hipGetLastError();
}

hipError_t launch_update_atlas_multi(const AtlasArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> items(G);
    for (int g = 0; g < G; g++) items[g] = uint32_t(a[g].n);
    return launch_per_item<AtlasArgs, k_update_atlas>(a, items.data(), G, ar, s, 256);
}

// flags[i] = 1 for leaves in use: mode 0 tests mFlags (atlas assignment),
// mode 1 tests mParent != ID_UNDEF64 (gvdbSetFlagSubcell,
// kernels/cuda_gvdb_particles.cuh:109-118).
__device__ inline void k_leaf_flags_body(const FlagArgs& a, uint32_t bid, uint32_t) {
    int i = bid * 256 + threadIdx.x;
    if (i >= a.n) return;
    const Node* n = a.t.nodes[0] + i;
    a.flags[i] = a.mode == 0 ? (n->mFlags ? 1u : 0u) : (n->mParent == ID_UNDEF64 ? 0u : 1u);
}
GVDB_MULTI(k_leaf_flags, FlagArgs, 256)

hipError_t launch_leaf_flags_multi(const FlagArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> items(G);
    for (int g = 0; g < G; g++) items[g] = uint32_t(a[g].n);
    return launch_per_item<FlagArgs, k_leaf_flags>(a, items.data(), G, ar, s, 256);
}

// -------------------------------------------------------- subcell binning
// w = p + trans.  BIN_FLOAT: count pass p.z != NOHIT, w >= 0; insert pass
// p >= 0 (gvdbInsertSubcell, kernels/cuda_gvdb_particles.cuh:218-273).
// BIN_FP16: insert pass w.z != NOHIT, w >= 0 (gvdbInsertSubcell_fp16,
// :154-216).  Non-finite positions are never binned.
__device__ inline bool binned_point(float p0, float p1, float p2, const float w[3], int rule) {
    if (p2 == NOHIT || !isfinite(w[0]) || !isfinite(w[1]) || !isfinite(w[2])) return false;
    if (w[0] < 0 || w[1] < 0 || w[2] < 0) return false;
    return rule == BIN_FP16 ? w[2] != NOHIT : (p0 >= 0 && p1 >= 0 && p2 >= 0);
}

// Each point is binned into every subcell overlapping
// [int(p-r)/sc*sc, int(p+1+r)/sc*sc] per axis whose brick exists
// (gvdbCountSubcell / gvdbInsertSubcell, kernels/cuda_gvdb_particles.cuh:218-315).
// The reference's count pass and its insert pass test different things (see
// BIN_FLOAT / BIN_FP16); where they disagree it writes past the counted
// entries, so a point is binned here only when both tests pass.
// The last leaf found is reused while the subcell corner stays inside it.
template <bool INSERT>
__global__ __launch_bounds__(256) void k_bin_subcell(Tree t, PointSrc pts, int scs, int spb, float tx, float ty, float tz,
                                                     float radius, const uint32_t* mapping, uint32_t* sc_cnt,
                                                     const uint32_t* sc_off, float* sc_pnt, int rule) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pts.n) return;
    const float* p = point_at(pts, i);
    const float p0 = p[0], p1 = p[1], p2 = p[2];
    float w[3] = {p0 + tx, p1 + ty, p2 + tz};
    if (!binned_point(p0, p1, p2, w, rule)) return;
    int scmin[3], scmax[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        scmin[a] = (int(w[a] - radius) / scs) * scs;
        scmax[a] = (int(w[a] + 1 + radius) / scs) * scs;
    }
    const int res0 = t.res[0];
    const int scdim = res0 / scs;
    int cache_leaf = -1, cx = 0, cy = 0, cz = 0;
    for (int sx = scmin[0]; sx <= scmax[0]; sx += scs)
        for (int sy = scmin[1]; sy <= scmax[1]; sy += scs)
            for (int sz = scmin[2]; sz <= scmax[2]; sz += scs) {
                int leaf;
                if (cache_leaf >= 0 && sx >= cx && sy >= cy && sz >= cz && sx < cx + res0 && sy < cy + res0 && sz < cz + res0) {
                    leaf = cache_leaf;
                } else {
                    leaf = pos_leaf(t, sx, sy, sz);
                    if (leaf < 0) continue;
                    const Node* ln = t.nodes[0] + leaf;
                    cache_leaf = leaf;
                    cx = ln->mPos[0]; cy = ln->mPos[1]; cz = ln->mPos[2];
                }
                // A subcell corner less than one cell below a node's corner
                // truncates into the node's first cell (the reference's walk),
                // and the local offset goes negative: the entry lands in a
                // neighbouring subcell, or before the first linked brick's
                // subcells, which no table holds (skipped here).
                // A top node that is itself a leaf is returned for every
                // position, inside it or not: corners past its far side have
                // no subcell either.
                int lx = (sx - cx) / scs, ly = (sy - cy) / scs, lz = (sz - cz) / scs;
                if (lx >= scdim || ly >= scdim || lz >= scdim) continue;
                const long long sidx = (long long)mapping[leaf] * spb + (lz * scdim + ly) * scdim + lx;
                if (sidx < 0) continue;
                const u64 idx = u64(sidx);
                uint32_t slot = atomicAdd(&sc_cnt[idx], 1u);
                if (INSERT) {
                    u64 e = u64(sc_off[idx]) + slot;
                    sc_pnt[3 * e] = w[0];
                    sc_pnt[3 * e + 1] = w[1];
                    sc_pnt[3 * e + 2] = w[2];
                }
            }
}

hipError_t launch_count_subcell(const Tree& t, PointSrc pts, int scs, int spb, const float trans[3], float radius,
                                const uint32_t* mapping, uint32_t* sc_cnt, int rule, hipStream_t s) {
    if (pts.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin_subcell<false>, dim3((pts.n + 255) / 256), dim3(256), 0, s, t, pts, scs, spb, trans[0],
                       trans[1], trans[2], radius, mapping, sc_cnt, (const uint32_t*)nullptr, (float*)nullptr, rule);
    return hipGetLastError();
}

hipError_t launch_insert_subcell(const Tree& t, PointSrc pts, int scs, int spb, const float trans[3], float radius,
                                 const uint32_t* mapping, uint32_t* sc_cnt, const uint32_t* sc_off, float* sc_pnt,
                                 int rule, hipStream_t s) {
    if (pts.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin_subcell<true>, dim3((pts.n + 255) / 256), dim3(256), 0, s, t, pts, scs, spb, trans[0],
                       trans[1], trans[2], radius, mapping, sc_cnt, sc_off, sc_pnt, rule);
    return hipGetLastError();
}

// gvdbCalcSubcellPos: kernels/cuda_gvdb_particles.cuh:677-694
__device__ inline void k_subcell_pos_body(const SubPosArgs& a, uint32_t bid, uint32_t) {
    const Tree& t = a.t;
    const int scs = a.scs, spb = a.spb;
    const uint32_t* mapping = a.mapping;
    int* sc_pos = a.pos;
    int* sc_nid = a.nid;
    int i = bid * 256 + threadIdx.x;
    if (i >= a.n) return;
    const Node* n = t.nodes[0] + i;
    if (n->mParent == ID_UNDEF64) return;
    int scdim = t.res[0] / scs;
    for (int k = 0; k < spb; k++) {
        u64 idx = u64(mapping[i]) * spb + k;
        sc_pos[3 * idx] = n->mPos[0] + k % scdim * scs;
        sc_pos[3 * idx + 1] = n->mPos[1] + (k / scdim) % scdim * scs;
        sc_pos[3 * idx + 2] = n->mPos[2] + (k / scdim / scdim) % scdim * scs;
        sc_nid[idx] = i;
    }
}

GVDB_MULTI(k_subcell_pos, SubPosArgs, 256)

hipError_t launch_subcell_pos_multi(const SubPosArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> items(G);
    for (int g = 0; g < G; g++) items[g] = uint32_t(a[g].n);
    return launch_per_item<SubPosArgs, k_subcell_pos>(a, items.data(), G, ar, s, 256);
}

// ---------------------------------------------------------- GatherDensity
// d^2 = |p - v|^2 as an unevaluated sum hi + lo, accurate to about 2^-44
// relative: the differences are split into float + error (TwoSum, exact),
// the squares into product + FMA residual, and the parts summed with TwoSum.
// The dropped terms (error squared, rounding of the small parts) stay below
// 2^-44 d^2.  The differences alone are exact in float only when p is not
// much smaller than v, which fails near the origin.
__device__ inline float two_diff(float a, float b, float& e) {
#pragma clang fp contract(off)
    const float s = a - b;
    const float bb = s - a;
    e = (a - (s - bb)) - (b + bb);
    return s;
}

// The least integer above the exact p - r: p - r = h + e exactly (TwoSum); an
// integer cannot lie strictly between h and h + e (it would be nearer), so
// floor(p - r) = floor(h), less one when h is an integer and e < 0.
__device__ inline int first_above(float p, float r) {
    float e;
    const float h = two_diff(p, r, e);
    int f = int(floorf(h));
    if (float(f) == h && e < 0.0f) f--;
    return f + 1;
}

__device__ inline void exact_d2(float px, float py, float pz, float vx, float vy, float vz, float& hi, float& lo) {
#pragma clang fp contract(off)
    float ex, ey, ez;
    const float dx = two_diff(px, vx, ex), dy = two_diff(py, vy, ey), dz = two_diff(pz, vz, ez);
    float hx = dx * dx, hy = dy * dy, hz = dz * dz;
    float lx = __builtin_fmaf(dx, dx, -hx), ly = __builtin_fmaf(dy, dy, -hy), lz = __builtin_fmaf(dz, dz, -hz);
    float s1 = hx + hy;
    float b1 = s1 - hx;
    float e1 = (hx - (s1 - b1)) + (hy - b1);
    float s2 = s1 + hz;
    float b2 = s2 - s1;
    float e2 = (s1 - (s2 - b2)) + (hz - b2);
    float cross = 2.0f * __builtin_fmaf(dx, ex, __builtin_fmaf(dy, ey, dz * ez));
    float tail = (e1 + e2) + ((lx + ly + lz) + cross);
    hi = s2 + tail;
    lo = tail - (hi - s2);
}

// radius - |p - v| from the accurate d^2: r^2 - d^2 divided by r + |p - v|,
// so the result keeps its relative accuracy down to the edge of the falloff.
__device__ inline float falloff_inl(float px, float py, float pz, int vx, int vy, int vz, float r) {
#pragma clang fp contract(off)
    float hi, lo;
    exact_d2(px, py, pz, float(vx), float(vy), float(vz), hi, lo);
    float r2h = r * r;
    float r2l = __builtin_fmaf(r, r, -r2h);
    float num = (r2h - hi) + (r2l - lo);
    float c = __builtin_sqrtf(hi);
    return num / (r + c);
}

__device__ __noinline__ float falloff_accurate(float px, float py, float pz, int vx, int vy, int vz, float r) {
    return falloff_inl(px, py, pz, vx, vy, vz, r);
}

// One workgroup per brick, one thread per interior voxel; the brick's
// subcell entries are staged in LDS and every voxel keeps the nearest entry
// of its subcell (gvdbGatherDensity, kernels/cuda_gvdb_particles.cuh:696-733:
// val = max(val, r - |p_j - v|), v the integer voxel coordinate).
static const int GATHER_CHUNK = 1024;

template <typename AT>
__device__ inline void store_atlas(AT* a, u64 off, float v);
template <>
__device__ inline void store_atlas<float>(float* a, u64 off, float v) { a[off] = v; }
template <>
__device__ inline void store_atlas<__half>(__half* a, u64 off, float v) { a[off] = __float2half(v); }
template <typename AT>
__device__ inline float load_atlas(const AT* a, u64 off);
template <>
__device__ inline float load_atlas<float>(const float* a, u64 off) { return a[off]; }
template <>
__device__ inline float load_atlas<__half>(const __half* a, u64 off) { return __half2float(a[off]); }

// gvdbGatherLevelSet_fp16 (kernels/cuda_gvdb_particles.cuh:776-811) reads a
// coordinate as u / 65535.0f * range + min; the product and the sum are
// fused, as a contracting compiler emits them.
__device__ inline float dequant16(uint16_t u, float mn, float rg) {
    return __fmaf_rn(__fdiv_rn(float(u), 65535.0f), rg, mn);
}

template <typename AT>
__device__ inline void k_gather_body(const GatherArgs& a, uint32_t bid, uint32_t) {
    const Tree& t = a.t;
    const int scs = a.scs;
    const float radius = a.radius;
    const int* sc_nid = a.sc_nid;
    const uint32_t *sc_cnt = a.sc_cnt, *sc_use = a.sc_use, *sc_off = a.sc_off;
    const int* sc_pos = a.sc_pos;
    const float* sc_pnt = a.sc_pnt;
    const uint16_t* sc_pnt16 = a.sc_pnt16;
    AT* atlas = reinterpret_cast<AT*>(a.atlas);
    const int accumulate = a.accumulate;
    __shared__ float lx[GATHER_CHUNK], ly[GATHER_CHUNK], lz[GATHER_CHUNK];
    const int res0 = t.res[0];
    const int scs3 = scs * scs * scs;
    const int spb = (res0 / scs) * (res0 / scs) * (res0 / scs);
    const int nthreads = res0 * res0 * res0;
    const u64 brick = bid;
    const int tid = threadIdx.x;
    const bool active = tid < nthreads;
    const int k = active ? tid / scs3 : 0;      // subcell within the brick
    const int q = tid - k * scs3;               // voxel within the subcell
    const u64 sc = brick * spb + k;
    int vx = 0, vy = 0, vz = 0;
    uint32_t beg = 0, cnt = 0;
    if (active) {
        vx = sc_pos[3 * sc] + q % scs;
        vy = sc_pos[3 * sc + 1] + (q / scs) % scs;
        vz = sc_pos[3 * sc + 2] + q / (scs * scs);
        beg = sc_off[sc];
        // the level set and accumulation take every entry; the density only
        // those that can raise a voxel (placed first)
        cnt = (accumulate || a.levelset) ? sc_cnt[sc] : sc_use[sc];
    }
    const u64 sc0 = brick * spb;
    const uint32_t bbeg = sc_off[sc0];
    const uint32_t bend = sc_off[sc0 + spb - 1] + sc_cnt[sc0 + spb - 1];
    const float fx = float(vx), fy = float(vy), fz = float(vz);
    float best = INFINITY;
    uint32_t arg = 0xFFFFFFFFu;
    for (uint32_t c0 = bbeg; c0 < bend; c0 += GATHER_CHUNK) {
        uint32_t c1 = min(bend, c0 + GATHER_CHUNK);
        __syncthreads();
        if (sc_pnt16) {
            for (uint32_t e = c0 + tid; e < c1; e += blockDim.x) {
                lx[e - c0] = dequant16(sc_pnt16[3 * u64(e)], a.qmin[0], a.qrange[0]);
                ly[e - c0] = dequant16(sc_pnt16[3 * u64(e) + 1], a.qmin[1], a.qrange[1]);
                lz[e - c0] = dequant16(sc_pnt16[3 * u64(e) + 2], a.qmin[2], a.qrange[2]);
            }
        } else {
            for (uint32_t e = c0 + tid; e < c1; e += blockDim.x) {
                lx[e - c0] = sc_pnt[3 * u64(e)];
                ly[e - c0] = sc_pnt[3 * u64(e) + 1];
                lz[e - c0] = sc_pnt[3 * u64(e) + 2];
            }
        }
        __syncthreads();
        uint32_t j0 = max(beg, c0), j1 = min(beg + cnt, c1);
        for (uint32_t j = j0; j < j1; j++) {
            float dx = lx[j - c0] - fx, dy = ly[j - c0] - fy, dz = lz[j - c0] - fz;
            float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
            if (d2 < best) { best = d2; arg = j; }
        }
    }
    if (!active) return;
    const Node* nd = t.nodes[0] + sc_nid[sc];
    if (!nd->mFlags || nd->mValue[0] < 0) return;   // no atlas slot to write
    int ax = nd->mValue[0] + (vx - nd->mPos[0]);
    int ay = nd->mValue[1] + (vy - nd->mPos[1]);
    int az = nd->mValue[2] + (vz - nd->mPos[2]);
    u64 off = atlas_offset(t, ax, ay, az);
    float val = accumulate ? load_atlas<AT>(atlas, off) : (a.levelset ? 3.0f : 0.0f);
    if (arg != 0xFFFFFFFFu) {
        float p[3];
        if (sc_pnt16) {
#pragma unroll
            for (int ax3 = 0; ax3 < 3; ax3++) p[ax3] = dequant16(sc_pnt16[3 * u64(arg) + ax3], a.qmin[ax3], a.qrange[ax3]);
        } else {
#pragma unroll
            for (int ax3 = 0; ax3 < 3; ax3++) p[ax3] = sc_pnt[3 * u64(arg) + ax3];
        }
        float f = falloff_accurate(p[0], p[1], p[2], vx, vy, vz, radius);
        val = a.levelset ? fminf(val, -f) : fmaxf(val, f);
    }
    store_atlas<AT>(atlas, off, val);
}

// Row-parallel gather for 8^3 bricks with 4^3 subcells: one wave per subcell,
// lane l owns the 4-voxel x-row (y, z) = (l & 3, (l >> 2) & 3) of the subcell
// for the entries j = l >> 4 (mod 4), so a wave step evaluates four entries on
// all 64 voxels; the four partial argmins per voxel are merged with shuffles
// (smallest d^2, ties to the earlier entry, as one sequential pass).
template <typename AT>
__device__ inline void k_gather_rows_body(const GatherArgs& a, uint32_t bid, uint32_t) {
    const Tree& t = a.t;
    const float radius = a.radius;
    const int* sc_nid = a.sc_nid;
    const uint32_t *sc_use = a.sc_use, *sc_off = a.sc_off;
    const int* sc_pos = a.sc_pos;
    const float* sc_pnt = a.sc_pnt;
    AT* atlas = reinterpret_cast<AT*>(a.atlas);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const u64 sc = u64(bid) * 8 + wave;
    const int row = lane & 15, g = lane >> 4;
    const int ry = row & 3, rz = row >> 2;
    const int sx = sc_pos[3 * sc], sy = sc_pos[3 * sc + 1], sz = sc_pos[3 * sc + 2];
    const uint32_t beg = sc_off[sc], use = sc_use[sc];
    const float fy = float(sy + ry), fz = float(sz + rz);
    const float fx0 = float(sx), fx1 = float(sx + 1), fx2 = float(sx + 2), fx3 = float(sx + 3);
    float b0 = INFINITY, b1 = INFINITY, b2 = INFINITY, b3 = INFINITY;
    uint32_t a0 = 0xFFFFFFFFu, a1 = 0xFFFFFFFFu, a2 = 0xFFFFFFFFu, a3 = 0xFFFFFFFFu;
    const float* base = sc_pnt + 3 * u64(beg);
    // 64 entries per chunk: one coalesced load per lane, broadcast by shuffles
    for (uint32_t c0 = 0; c0 < use; c0 += 64) {
        const uint32_t jl = c0 + lane;
        float qx = 0.f, qy = 0.f, qz = 0.f;
        if (jl < use) {
            const float* q = base + 3 * u64(jl);
            qx = q[0]; qy = q[1]; qz = q[2];
        }
        const uint32_t nn = min(64u, use - c0);
        for (uint32_t k = 0; k < nn; k += 4) {
            const int src = int(k) + g;
            const float px = __shfl(qx, src, 64), py = __shfl(qy, src, 64), pz = __shfl(qz, src, 64);
            if (uint32_t(src) < nn) {
                const uint32_t j = c0 + uint32_t(src);
                const float dy = py - fy, dz = pz - fz;
                const float dyz = __builtin_fmaf(dy, dy, dz * dz);
                float dx, d2;
                dx = px - fx0; d2 = __builtin_fmaf(dx, dx, dyz); if (d2 < b0) { b0 = d2; a0 = j; }
                dx = px - fx1; d2 = __builtin_fmaf(dx, dx, dyz); if (d2 < b1) { b1 = d2; a1 = j; }
                dx = px - fx2; d2 = __builtin_fmaf(dx, dx, dyz); if (d2 < b2) { b2 = d2; a2 = j; }
                dx = px - fx3; d2 = __builtin_fmaf(dx, dx, dyz); if (d2 < b3) { b3 = d2; a3 = j; }
            }
        }
    }
    float bb[4] = {b0, b1, b2, b3};
    uint32_t aa[4] = {a0, a1, a2, a3};
#pragma unroll
    for (int sh = 16; sh <= 32; sh <<= 1) {
#pragma unroll
        for (int x = 0; x < 4; x++) {
            float ob = __shfl_xor(bb[x], sh, 64);
            uint32_t oa = __shfl_xor(aa[x], sh, 64);
            if (ob < bb[x] || (ob == bb[x] && oa < aa[x])) { bb[x] = ob; aa[x] = oa; }
        }
    }
    if (g != 0) return;
    const Node* nd = t.nodes[0] + sc_nid[sc];
    if (!nd->mFlags || nd->mValue[0] < 0) return;   // no atlas slot to write
    const int vy = sy + ry, vz = sz + rz;
    const u64 off = atlas_offset(t, nd->mValue[0] + (sx - nd->mPos[0]), nd->mValue[1] + (vy - nd->mPos[1]),
                                 nd->mValue[2] + (vz - nd->mPos[2]));
#pragma unroll
    for (int x = 0; x < 4; x++) {
        float val = 0.0f;
        if (aa[x] != 0xFFFFFFFFu) {
            const float* p = base + 3 * u64(aa[x]);
            val = fmaxf(0.0f, falloff_accurate(p[0], p[1], p[2], sx + x, vy, vz, radius));
        }
        store_atlas<AT>(atlas, off + x, val);
    }
}

GVDB_MULTI_T(k_gather, typename AT, AT, GatherArgs, 512)

// gvdbInsertSubcell_fp16 (kernels/cuda_gvdb_particles.cuh:204-207):
// u16((w - pos_min) / pos_range * 65535), truncated; values outside
// [0, 65535] (points outside the bounds RebuildTopology measured) saturate.
__global__ __launch_bounds__(256) void k_quantize(const float* src, uint16_t* dst, u64 n3, float m0, float m1, float m2,
                                                  float r0, float r1, float r2) {
    const u64 i = u64(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n3) return;
    const int ax = int(i % 3);
    const float mn = ax == 0 ? m0 : (ax == 1 ? m1 : m2);
    const float rg = ax == 0 ? r0 : (ax == 1 ? r1 : r2);
    const float q = __fmul_rn(__fdiv_rn(__fsub_rn(src[i], mn), rg), 65535.0f);
    dst[i] = q >= 65535.0f ? uint16_t(65535) : (q > 0.0f ? uint16_t(int(q)) : uint16_t(0));
}

hipError_t launch_quantize_entries(const float* src, uint16_t* dst, u64 n3, const float qmin[3], const float qrange[3],
                                   hipStream_t s) {
    if (n3 == 0) return hipSuccess;
    if (!src || !dst) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_quantize, dim3(unsigned((n3 + 255) / 256)), dim3(256), 0, s, src, dst, n3, qmin[0], qmin[1],
                       qmin[2], qrange[0], qrange[1], qrange[2]);
    return hipGetLastError();
}
GVDB_MULTI_T(k_gather_rows, typename AT, AT, GatherArgs, 512)

// One launch per (kernel, channel type) group: 8^3 bricks with 4^3 subcells
// and no accumulation take the row-parallel kernel.
hipError_t launch_gather_multi(const GatherArgs* a, int G, Arena* ar, hipStream_t s) {
    for (int variant = 0; variant < 4; variant++) {
        const bool rows = variant & 1, half = variant >> 1;
        std::vector<GatherArgs> sel;
        std::vector<uint32_t> nb;
        for (int g = 0; g < G; g++) {
            const bool r = !a[g].accumulate && !a[g].levelset && !a[g].sc_pnt16 && a[g].t.res[0] == 8 && a[g].scs == 4;
            if (r != rows || (a[g].atlas_type == T_HALF) != half || a[g].nbricks <= 0) continue;
            if (a[g].t.res[0] * a[g].t.res[0] * a[g].t.res[0] > 512) return hipErrorInvalidValue;
            if (!a[g].sc_nid || !a[g].sc_cnt || !a[g].sc_use || !a[g].sc_off || !a[g].sc_pos || !a[g].atlas ||
                (a[g].sc_pnt16 && !a[g].levelset))
                return hipErrorInvalidValue;
            sel.push_back(a[g]);
            nb.push_back(uint32_t(a[g].nbricks));
        }
        if (sel.empty()) continue;
        MultiLaunch L;
        hipError_t e = multi_setup(ar, sel.data(), nb.data(), int(sel.size()), s, L);
        if (e != hipSuccess) return e;
        if (L.blocks == 0) continue;
        if (rows && half) hipLaunchKernelGGL(k_gather_rows<__half>, dim3(L.blocks), dim3(512), 0, s, sel[0], L.map);
        else if (rows) hipLaunchKernelGGL(k_gather_rows<float>, dim3(L.blocks), dim3(512), 0, s, sel[0], L.map);
        else if (half) hipLaunchKernelGGL(k_gather<__half>, dim3(L.blocks), dim3(512), 0, s, sel[0], L.map);
        else hipLaunchKernelGGL(k_gather<float>, dim3(L.blocks), dim3(512), 0, s, sel[0], L.map);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
}  // namespace gvdb

namespace gvdb {

// ComputeBounds (src/gvdb_volume_gvdb.cpp:1792-1816): min corner and max
// corner + range of the flagged leaves.  out6 starts as {INT_MAX x3, INT_MIN x3}.
__device__ inline void k_leaf_bounds_body(const BoundsArgs& ba, uint32_t bid, uint32_t nblk) {
    const Tree& t = ba.t;
    int* out6 = ba.out6;
    int mn[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF}, mx[3] = {(int)0x80000000, (int)0x80000000, (int)0x80000000};
    for (int i = bid * 256 + threadIdx.x; i < ba.n; i += nblk * 256) {
        const Node* n = t.nodes[0] + i;
        if (!n->mFlags) continue;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = min(mn[a], n->mPos[a]);
            mx[a] = max(mx[a], n->mPos[a] + t.range[0]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            mn[a] = min(mn[a], __shfl_xor(mn[a], d, 64));
            mx[a] = max(mx[a], __shfl_xor(mx[a], d, 64));
        }
    }
    __shared__ int smn[4][3], smx[4][3];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; a++) { smn[wave][a] = mn[a]; smx[wave][a] = mx[a]; }
    __syncthreads();
    if (threadIdx.x < 3) {
        int a = threadIdx.x, lo = smn[0][a], hi = smx[0][a];
        for (int w = 1; w < 4; w++) { lo = min(lo, smn[w][a]); hi = max(hi, smx[w][a]); }
        atomicMin(&out6[a], lo);
        atomicMax(&out6[3 + a], hi);
    }
}

GVDB_MULTI(k_leaf_bounds, BoundsArgs, 256)

hipError_t launch_leaf_bounds_multi(const BoundsArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<FillArgs> f;
    std::vector<uint32_t> nb(G);
    for (int g = 0; g < G; g++) {
        f.push_back(FillArgs{a[g].out6, 3, 0x7FFFFFFFull, 0});
        f.push_back(FillArgs{a[g].out6 + 3, 3, 0x80000000ull, 0});
        uint32_t b = uint32_t((a[g].n + 255) / 256);
        nb[g] = b > 256 ? 256 : b;
    }
    hipError_t e = launch_fill_multi(f.data(), int(f.size()), ar, s);
    if (e != hipSuccess) return e;
    MultiLaunch L;
    e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    hipLaunchKernelGGL(k_leaf_bounds, dim3(L.blocks), dim3(256), 0, s, a[0], L.map);
    return hipGetLastError();
}

}  // namespace gvdb

// ===================================================== pull-based binning
// The reference bins by pushing every (point, subcell) pair through a global
// atomic counter, twice (count, then insert).  Here each point is bucketed
// once by the subcell that contains it, and every brick then pulls the points
// of the subcells around it: a point can only be binned into subcells at most
// R = ceil((r+1)/sc) subcells away from its own, so a brick reads the
// (spd+2R)^3 subcells around its spd^3 own ones and applies the reference's
// per-axis test  int(p-r)/sc*sc <= s <= int(p+1+r)/sc*sc  to each.  Counts
// and placement are per workgroup (LDS and wave ballots), without global
// atomics.  Used when every valid point lies in an existing brick and the
// neighbourhood fits in the 27 bricks around; otherwise the push kernels run.
namespace gvdb {

static const int LEN_PARTS = 64;   // spread counters of the entry count

__device__ inline void k_bucket_points_body(const BucketArgs& a, uint32_t bid, uint32_t) {
    const Tree& t = a.t;
    const PointSrc& pts = a.pts;
    const int scs = a.scs, spb = a.spb;
    const float tx = a.tr[0], ty = a.tr[1], tz = a.tr[2];
    uint32_t *key = a.key, *slot = a.slot, *bcnt = a.bcnt, *orphans = a.orphans;
    const int i = bid * 256 + threadIdx.x;
    uint32_t k = 0xFFFFFFFFu, binned = 0;
    if (i < pts.n) {
        const float* p = point_at(pts, i);
        float p0 = p[0], p1 = p[1], p2 = p[2];
        float w[3] = {p0 + tx, p1 + ty, p2 + tz};
        const bool valid = binned_point(p0, p1, p2, w, a.rule);
        if (valid) {
            // the subcell size divides the (power-of-two) brick width, so it is
            // a power of two: divisions are shifts (towards zero for negatives)
            const int ssh = __builtin_ctz(unsigned(scs));
            int ox = (int(w[0]) >> ssh) << ssh, oy = (int(w[1]) >> ssh) << ssh, oz = (int(w[2]) >> ssh) << ssh;
            int leaf, org[3];
            uint32_t tmask = 0;
            if (a.bt.tab) {
                // the brick table: the leaf whose box holds the subcell corner,
                // and its neighbour mask, in one read
                const int bsh = __builtin_ctz(unsigned(t.res[0]));
                const unsigned x = unsigned((ox >> bsh) - a.bt.lo[0]), y = unsigned((oy >> bsh) - a.bt.lo[1]),
                               z = unsigned((oz >> bsh) - a.bt.lo[2]);
                leaf = -1;
                if (x < unsigned(a.bt.dim[0]) && y < unsigned(a.bt.dim[1]) && z < unsigned(a.bt.dim[2])) {
                    const u64 e = a.bt.tab[tab_index(a.bt, x, y, z)];
                    if (uint32_t(e) != ~0u) leaf = int(uint32_t(e));
                    tmask = uint32_t(e >> 32);
                }
                org[0] = (ox >> bsh) << bsh, org[1] = (oy >> bsh) << bsh, org[2] = (oz >> bsh) << bsh;
            } else {
                leaf = pos_leaf(t, ox, oy, oz);
                if (leaf >= 0) {
                    const Node* n = t.nodes[0] + leaf;
                    org[0] = n->mPos[0], org[1] = n->mPos[1], org[2] = n->mPos[2];
                    // a subcell below the found brick's corner (the walk truncates
                    // positions less than a cell below a node into its first cell),
                    // or past a top node that is a leaf, has no bucket: such points
                    // go to the push path with the orphans
                    const int res0 = t.res[0];
                    if (ox < org[0] || oy < org[1] || oz < org[2] || ox >= org[0] + res0 || oy >= org[1] + res0 ||
                        oz >= org[2] + res0)
                        leaf = -1;
                }
            }
            if (leaf < 0) {
                atomicAdd(orphans, 1u);
            } else {
                const int res0 = t.res[0];
                int sd = res0 >> ssh;
                int lx = (ox - org[0]) >> ssh, ly = (oy - org[1]) >> ssh, lz = (oz - org[2]) >> ssh;
                k = uint32_t(leaf) * uint32_t(spb) + uint32_t((lz * sd + ly) * sd + lx);
                slot[i] = atomicAdd(&bcnt[k], 1u);
                if (a.len) {
                    // subcells the point is binned into (CountSubcell's per-axis
                    // range), split per axis into those below, inside and above
                    // this brick; all lie in it or a face/edge/corner neighbour
                    uint32_t c[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                    for (int ax = 0; ax < 3; ax++) {
                        const int lo = child_cell(int(w[ax] - a.radius), ssh) * scs,
                                  hi = child_cell(int(w[ax] + 1 + a.radius), ssh) * scs;
                        for (int sv = lo; sv <= hi; sv += scs) c[ax][sv < org[ax] ? 0 : (sv >= org[ax] + res0 ? 2 : 1)]++;
                    }
                    const uint32_t m = a.bt.tab ? tmask : a.nmask[leaf];
#pragma unroll
                    for (int d = 0; d < 27; d++)
                        if ((m >> d) & 1u) binned += c[0][d % 3] * c[1][(d / 3) % 3] * c[2][d / 9];
                }
            }
        }
        key[i] = k;
    }
    if (a.len) {
        // block total, added to one of LEN_PARTS counters (one address would
        // serialise every block's atomic)
        __shared__ uint32_t part[4];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) binned += __shfl_xor(binned, d, 64);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = binned;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long sum = u64(part[0]) + part[1] + part[2] + part[3];
            if (sum) atomicAdd(reinterpret_cast<unsigned long long*>(a.len) + (bid % LEN_PARTS), sum);
        }
    }
}

// Sum of a grid's LEN_PARTS entry counters into its result word.
struct SumArgs {
    const u64* parts;
    u64* out;
};
__device__ inline void k_sum_parts_body(const SumArgs& a, uint32_t, uint32_t) {
    u64 v = threadIdx.x < LEN_PARTS ? a.parts[threadIdx.x] : 0ull;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (threadIdx.x == 0) *a.out = v;
}
GVDB_MULTI(k_sum_parts, SumArgs, 64)

hipError_t launch_sum_parts_multi(const u64* const* parts, u64* const* out, int G, Arena* ar, hipStream_t s) {
    std::vector<SumArgs> a(G);
    std::vector<uint32_t> nb(G, 1u);
    for (int g = 0; g < G; g++) a[g] = SumArgs{parts[g], out[g]};
    MultiLaunch L;
    hipError_t e = multi_setup(ar, a.data(), nb.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    hipLaunchKernelGGL(k_sum_parts, dim3(L.blocks), dim3(64), 0, s, a[0], L.map);
    return hipGetLastError();
}

__device__ inline void k_bucket_scatter_body(const ScatterArgs& a, uint32_t bid, uint32_t) {
    const PointSrc& pts = a.pts;
    const float tx = a.tr[0], ty = a.tr[1], tz = a.tr[2];
    const uint32_t *key = a.key, *slot = a.slot, *boff = a.boff;
    float4* sorted = a.sorted;
    int i = bid * 256 + threadIdx.x;
    if (i >= pts.n) return;
    uint32_t k = key[i];
    if (k == 0xFFFFFFFFu) return;
    const float* p = point_at(pts, i);
    u64 e = u64(boff[k]) + slot[i];
    const uint32_t c = a.clr.base ? *reinterpret_cast<const uint32_t*>(a.clr.base + u64(i) * a.clr.stride + a.clr.off) : 0u;
    sorted[e] = make_float4(p[0] + tx, p[1] + ty, p[2] + tz, __uint_as_float(c));
}

static const int PULL_THREADS = 256;
static const int PULL_MAX_CELLS = 512;   // (spd + 2R)^3 upper bound handled
static const int PULL_MAX_SPB = 32;
static const int PULL_CAP = 4096;        // pulled points per LDS window

struct PullShared {
    uint32_t cell_off[PULL_MAX_CELLS];
    uint32_t cell_start[PULL_MAX_CELLS];
    uint16_t owner[PULL_CAP];
    int nbr[27];
    uint32_t wsum[PULL_THREADS / 64];
    uint32_t wave_cnt[PULL_THREADS / 64][PULL_MAX_SPB];   // binned entries that can raise a voxel
    uint32_t wave_zro[PULL_THREADS / 64][PULL_MAX_SPB];   // binned entries that cannot
    uint32_t run[PULL_MAX_SPB];
    uint32_t runz[PULL_MAX_SPB];
    uint32_t use[PULL_MAX_SPB];
};

// Per-brick setup shared by the count and write passes: the bucket ranges of
// the (sd+2R)^3 subcells around the brick and their exclusive prefix
// (workgroup scan).  nbr27[13] < 0 marks a brick the top node cannot reach,
// which receives no entries.  Returns the number of points to test.
__device__ inline uint32_t pull_setup(const Tree& t, int brick, int scs, int R, const int* nbr27,
                                      const uint32_t* bcnt, const uint32_t* boff, PullShared& S) {
    const int sd = t.res[0] / scs;
    const int span = sd + 2 * R;
    const int ncell = span * span * span;
    const int spb = sd * sd * sd;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x < 27) S.nbr[threadIdx.x] = nbr27[u64(brick) * 27 + threadIdx.x];
    __syncthreads();
    const bool ok = S.nbr[13] >= 0;
    // each thread owns two consecutive cells for the scan
    uint32_t cnt2[2];
    for (int k = 0; k < 2; k++) {
        int c = 2 * threadIdx.x + k;
        uint32_t cnt = 0, off = 0;
        if (c < ncell && ok) {
            int cx = c % span - R, cy = (c / span) % span - R, cz = c / (span * span) - R;
            int dx = cx < 0 ? -1 : (cx >= sd ? 1 : 0), dy = cy < 0 ? -1 : (cy >= sd ? 1 : 0), dz = cz < 0 ? -1 : (cz >= sd ? 1 : 0);
            int nb = S.nbr[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
            if (nb >= 0) {
                int lx = cx - dx * sd, ly = cy - dy * sd, lz = cz - dz * sd;
                u64 b = u64(nb) * spb + u64((lz * sd + ly) * sd + lx);
                cnt = bcnt[b];
                off = boff[b];
            }
        }
        if (c < ncell) S.cell_off[c] = off;
        cnt2[k] = cnt;
    }
    uint32_t mine = cnt2[0] + cnt2[1];
    uint32_t x = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) S.wsum[wave] = x;
    __syncthreads();
    uint32_t pre = x - mine, total = 0;
    for (int w = 0; w < PULL_THREADS / 64; w++) {
        if (w < wave) pre += S.wsum[w];
        total += S.wsum[w];
    }
    int c0 = 2 * threadIdx.x;
    if (c0 < ncell) S.cell_start[c0] = pre;
    if (c0 + 1 < ncell) S.cell_start[c0 + 1] = pre + cnt2[0];
    return total;
}

// Cell of every pulled point in the window [w0, w0 + PULL_CAP), written by
// the cell owners (replaces a search per point).
__device__ inline void pull_owners(PullShared& S, int ncell, uint32_t w0, uint32_t T) {
    for (int c = threadIdx.x; c < ncell; c += PULL_THREADS) {
        uint32_t beg = S.cell_start[c];
        uint32_t end = (c + 1 < ncell) ? S.cell_start[c + 1] : T;
        uint32_t lo = max(beg, w0);
        for (uint32_t f = lo; f < end && f < w0 + PULL_CAP; f++) S.owner[f - w0] = uint16_t(c);
    }
}

// Integer interval [a, b] of v with |p - v| < r.  p - v is exact in float for
// these magnitudes, so the end tests are exact.
__device__ inline void hit_interval(float p, float r, int& a, int& b) {
    a = int(floorf(p - r));
    if (p - float(a) >= r) a++;
    if (p - float(a - 1) < r) a--;
    b = int(ceilf(p + r));
    if (float(b) - p >= r) b--;
    if (float(b + 1) - p < r) b++;
}

// Masks of one pulled point over the brick's spb subcells: `m` the subcells
// it is binned into (the reference's per-axis test), `u` those where it can
// raise some voxel (|p - v| < r on every axis for a voxel v of the subcell).
__device__ inline void pull_mask(const Node* brick, int scs, int sd, float r, float w0, float w1, float w2,
                                 uint32_t& m, uint32_t& u) {
    const float w[3] = {w0, w1, w2};
    uint32_t mm[3], uu[3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        int lo = (int(w[ax] - r) / scs) * scs, hi = (int(w[ax] + 1 + r) / scs) * scs;
        int a, b;
        hit_interval(w[ax], r, a, b);
        uint32_t mb = 0, ub = 0;
        for (int i = 0; i < sd; i++) {
            int s = brick->mPos[ax] + i * scs;
            mb |= (s >= lo && s <= hi) ? (1u << i) : 0u;
            ub |= (max(a, s) <= min(b, s + scs - 1)) ? (1u << i) : 0u;
        }
        mm[ax] = mb;
        uu[ax] = ub;
    }
    m = 0;
    u = 0;
    for (int z = 0; z < sd; z++)
        for (int y = 0; y < sd; y++)
            for (int x = 0; x < sd; x++) {
                uint32_t bit = 1u << ((z * sd + y) * sd + x);
                if ((mm[2] >> z) & (mm[1] >> y) & (mm[0] >> x) & 1u) m |= bit;
                if ((uu[2] >> z) & (uu[1] >> y) & (uu[0] >> x) & 1u) u |= bit;
            }
    u &= m;
}

// Entries inside a subcell are placed with the ones that can raise a voxel
// first (the reference leaves the order inside a subcell unspecified); their
// number per subcell goes to sc_use, so the gather can stop there.
template <bool WRITE>
__device__ inline void k_pull_body(const PullArgs& a, uint32_t bid, uint32_t) {
    const Tree& t = a.t;
    const int scs = a.scs, R = a.R;
    const float radius = a.radius;
    const int* nbr27 = a.nbr;
    const uint32_t *bcnt = a.bcnt, *boff = a.boff, *mapping = a.mapping, *sc_off = a.sc_off;
    const float4* sorted = a.sorted;
    uint32_t *sc_cnt = a.sc_cnt, *sc_use = a.sc_use;
    float* sc_pnt = a.sc_pnt;
    __shared__ PullShared S;
    const int brick = int(bid);
    const Node* bn = t.nodes[0] + brick;
    const int sd = t.res[0] / scs;
    const int spb = sd * sd * sd;
    const int span = sd + 2 * R;
    const int ncell = span * span * span;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool linked = bn->mParent != ID_UNDEF64;
    const u64 sc0 = linked ? u64(mapping[brick]) * spb : 0;
    if (threadIdx.x < PULL_MAX_SPB) {
        S.run[threadIdx.x] = 0;
        S.runz[threadIdx.x] = 0;
        S.use[threadIdx.x] = (WRITE && linked && threadIdx.x < spb) ? sc_use[sc0 + threadIdx.x] : 0u;
    }
    const uint32_t T = pull_setup(t, brick, scs, R, nbr27, bcnt, boff, S);
    __syncthreads();
    for (uint32_t w0 = 0; w0 < T; w0 += PULL_CAP) {
        pull_owners(S, ncell, w0, T);
        __syncthreads();
        const uint32_t wend = min(T, w0 + PULL_CAP);
        for (uint32_t base = w0; base < wend; base += PULL_THREADS) {
            const uint32_t f = base + threadIdx.x;
            uint32_t m = 0, u = 0;
            float w0f = 0, w1f = 0, w2f = 0;
            if (f < wend) {
                const int c = S.owner[f - w0];
                const u64 e = u64(S.cell_off[c]) + (f - S.cell_start[c]);
                const float4 q = sorted[e];
                w0f = q.x; w1f = q.y; w2f = q.z;
                pull_mask(bn, scs, sd, radius, w0f, w1f, w2f, m, u);
            }
            const uint32_t z = m & ~u;
            for (int k = 0; k < spb; k++) {
                u64 bu = __ballot((u >> k) & 1u);
                u64 bz = __ballot((z >> k) & 1u);
                if (lane == 0) {
                    S.wave_cnt[wave][k] = uint32_t(__popcll(bu));
                    S.wave_zro[wave][k] = uint32_t(__popcll(bz));
                }
            }
            __syncthreads();
            if (WRITE) {
                // exclusive rank = entries of earlier chunks + earlier waves + lower lanes
                const u64 lower = lane ? ((~0ull) >> (64 - lane)) : 0ull;
                for (int k = 0; k < spb; k++) {
                    bool onu = (u >> k) & 1u, onz = (z >> k) & 1u;
                    u64 bu = __ballot(onu);
                    u64 bz = __ballot(onz);
                    if (onu || onz) {
                        uint32_t before = 0;
                        for (int w = 0; w < wave; w++) before += onu ? S.wave_cnt[w][k] : S.wave_zro[w][k];
                        uint32_t rank = onu ? S.run[k] + before + uint32_t(__popcll(bu & lower))
                                            : S.use[k] + S.runz[k] + before + uint32_t(__popcll(bz & lower));
                        u64 e = u64(sc_off[sc0 + k]) + rank;
                        sc_pnt[3 * e] = w0f;
                        sc_pnt[3 * e + 1] = w1f;
                        sc_pnt[3 * e + 2] = w2f;
                    }
                }
            }
            __syncthreads();
            if (threadIdx.x < spb) {
                uint32_t su = 0, sz = 0;
                for (int w = 0; w < PULL_THREADS / 64; w++) { su += S.wave_cnt[w][threadIdx.x]; sz += S.wave_zro[w][threadIdx.x]; }
                S.run[threadIdx.x] += su;
                S.runz[threadIdx.x] += sz;
            }
            __syncthreads();
        }
    }
    if (!WRITE && linked && threadIdx.x < spb) {
        sc_cnt[sc0 + threadIdx.x] = S.run[threadIdx.x] + S.runz[threadIdx.x];
        sc_use[sc0 + threadIdx.x] = S.run[threadIdx.x];
    }
}

// ---------------------------------------------------- gather from buckets
// GatherDensity without materialised subcell lists.  A voxel's value is the
// falloff of the nearest point among those binned into its subcell; every
// point that can raise the voxel (|p - v| < r) is binned into it, and any
// other point gives r - |p - v| <= 0, so the nearest among all points around
// the subcell gives the same value.  One wave per 4^3 subcell, eight per
// 8^3 brick, with no workgroup barriers: the wave walks the buckets of the
// (2R+1)^3 subcells around its own, keeps the points whose box reaches it
// (staged in its own LDS list), and evaluates them row-parallel (lane =
// 4-voxel x-row x 4 candidates per step).  The nearest point is decided on
// float d^2, and whenever two candidates come within 2^-20 relative of each
// other on the exact d^2 (FMA products + TwoSum), so the winner is the truly
// nearest point (ties to the lower bucket position); its falloff is exact.
static const int GB_THREADS = 512;
static const int GB_CHUNK = 1024;   // candidates staged per pass (two per thread)

// Is candidate a (index ia) nearer to v than b (index ib)?  Accurate d^2, ties to the lower index.
__device__ __noinline__ bool nearer_exact(float ax, float ay, float az, uint32_t ia, float bx, float by, float bz,
                                          uint32_t ib, float vx, float vy, float vz) {
    float ha, la, hb, lb;
    exact_d2(ax, ay, az, vx, vy, vz, ha, la);
    exact_d2(bx, by, bz, vx, vy, vz, hb, lb);
    if (ha != hb) return ha < hb;
    if (la != lb) return la < lb;
    return ia < ib;
}

// Nearest candidate so far for one voxel: its bucket position (which also
// fetches its coordinates when an exact comparison or the final falloff needs
// them) and its float d^2 raised by NEAR_ULPS ulps.  Float d^2 (rounded
// differences, two FMAs) is within 3.5 ulps of the exact value, so a candidate
// below the raised bound may be nearer; one also below the lowered bound
// surely is, and the few in between are settled on the exact d^2.
static const uint32_t NEAR_ULPS = 8;

struct Best {
    float bt;       // d^2 + NEAR_ULPS ulps (INFINITY while empty)
    uint32_t id;
};

__device__ inline void best_offer(Best& B, float d2, float px, float py, float pz, uint32_t id, float vx, float vy,
                                  float vz, const float4* pts) {
    if (!(d2 < B.bt)) return;
    bool take = B.id == 0xFFFFFFFFu;
    if (!take) {
        const float b = __uint_as_float(__float_as_uint(B.bt) - NEAR_ULPS);
        take = d2 < __uint_as_float(__float_as_uint(b) - NEAR_ULPS) && b > 0.0f;
        if (!take) {
            const float4 q = pts[B.id];
            take = nearer_exact(px, py, pz, id, q.x, q.y, q.z, B.id, vx, vy, vz);
        }
    }
    if (take) {
        B.bt = __uint_as_float(__float_as_uint(d2) + NEAR_ULPS);
        B.id = id;
    }
}

// Cell ranges of the (2 + 2R)^3 subcells around a brick and their exclusive
// prefix (one cell per thread); returns the candidate count, 0 when the
// brick is unreachable.
struct GBShared {
    uint32_t cell_off[GB_THREADS];
    uint32_t cell_start[GB_THREADS];
    uint16_t owner[PULL_CAP];
    int nbr[27];
    uint32_t wsum[GB_THREADS / 64];
    float4 pt[GB_CHUNK];      // x, y, z, bucket position (bits)
    uint8_t um[GB_CHUNK];
    uint16_t lst[GB_THREADS / 64][GB_CHUNK];
};

// The brick's 27 neighbour ids, one per thread below 27 (loaded by the
// caller ahead of its own checks, so the load overlaps them).
__device__ inline int gb_nbr(const int* nbr27, int brick) {
    return threadIdx.x < 27 ? nbr27[u64(brick) * 27 + threadIdx.x] : 0;
}

template <int NT = GB_THREADS, class SH>
__device__ inline uint32_t gb_setup(int R, int nb_mine, const uint32_t* bcnt, const uint32_t* boff, SH& S) {
    const int sd = 2, span = sd + 2 * R, ncell = span * span * span;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x < 27) S.nbr[threadIdx.x] = nb_mine;
    __syncthreads();
    const bool ok = S.nbr[13] >= 0;
    const int c = threadIdx.x;
    uint32_t cnt = 0, off = 0;
    if (c < ncell && ok) {
        int cx = c % span - R, cy = (c / span) % span - R, cz = c / (span * span) - R;
        int dx = cx < 0 ? -1 : (cx >= sd ? 1 : 0), dy = cy < 0 ? -1 : (cy >= sd ? 1 : 0), dz = cz < 0 ? -1 : (cz >= sd ? 1 : 0);
        int nb = S.nbr[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
        if (nb >= 0) {
            int lx = cx - dx * sd, ly = cy - dy * sd, lz = cz - dz * sd;
            u64 b = u64(nb) * 8 + u64((lz * sd + ly) * sd + lx);
            cnt = bcnt[b];
            off = boff[b];
        }
    }
    if (c < ncell) S.cell_off[c] = off;
    uint32_t x = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) S.wsum[wave] = x;
    __syncthreads();
    uint32_t pre = x - cnt, total = 0;
    for (int w = 0; w < NT / 64; w++) {
        if (w < wave) pre += S.wsum[w];
        total += S.wsum[w];
    }
    if (c < ncell) S.cell_start[c] = pre;
    return total;
}

template <typename AT>
__device__ inline void k_bucket_gather_body(const BucketGatherArgs& a, uint32_t bid, uint32_t) {
    __shared__ GBShared S;
    const Tree& t = a.t;
    const int brick = int(bid);
    const Node* bn = t.nodes[0] + brick;
    if (bn->mParent == ID_UNDEF64) return;        // unlinked leaves have no subcells
    if (!bn->mFlags || bn->mValue[0] < 0) return;  // no atlas slot to write (uniform per brick)
    const float r = a.radius;
    const int R = a.R, span = 2 + 2 * R, ncell = span * span * span;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = lane & 15, g = lane >> 4;
    const int ry = row & 3, rz = row >> 2;
    // this wave's subcell (x fastest, as CalcSubcellPos numbers them)
    const int sx = bn->mPos[0] + (wave & 1) * 4, sy = bn->mPos[1] + ((wave >> 1) & 1) * 4,
              sz = bn->mPos[2] + (wave >> 2) * 4;
    const float fx[4] = {float(sx), float(sx + 1), float(sx + 2), float(sx + 3)};
    const float fy = float(sy + ry), fz = float(sz + rz);
    Best B[4];
#pragma unroll
    for (int x = 0; x < 4; x++) B[x] = Best{INFINITY, 0xFFFFFFFFu};
    const uint32_t T = gb_setup(R, gb_nbr(a.nbr, brick), a.bcnt, a.boff, S);
    __syncthreads();
    const int org[3] = {bn->mPos[0], bn->mPos[1], bn->mPos[2]};
    const u64 lower = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    for (uint32_t w0 = 0; w0 < T; w0 += PULL_CAP) {
        // cell of every candidate in the window
        for (int c = threadIdx.x; c < ncell; c += GB_THREADS) {
            uint32_t beg = S.cell_start[c];
            uint32_t end = (c + 1 < ncell) ? S.cell_start[c + 1] : T;
            for (uint32_t f = max(beg, w0); f < end && f < w0 + PULL_CAP; f++) S.owner[f - w0] = uint16_t(c);
        }
        __syncthreads();
        const uint32_t wend = min(T, w0 + PULL_CAP);
        for (uint32_t c0 = w0; c0 < wend; c0 += GB_CHUNK) {
            // candidates of this chunk: positions and the subcells of the brick each can reach
            for (int q = threadIdx.x; q < GB_CHUNK; q += GB_THREADS) {
                const uint32_t f = c0 + q;
                uint32_t um = 0;
                if (f < wend) {
                    const int c = S.owner[f - w0];
                    const uint32_t pos = S.cell_off[c] + (f - S.cell_start[c]);
                    const float4 p = a.sorted[pos];
                    const float pp[3] = {p.x, p.y, p.z};
                    uint32_t am[3];
#pragma unroll
                    for (int ax = 0; ax < 3; ax++) {
                        // the entry-list test of CountSubcell/InsertSubcell: subcell
                        // origin s is listed when int(p - r)/4*4 <= s <= int(p + 1 + r)/4*4
                        // (C truncations), p including trans as bucketed
                        const int lo = child_cell(int(pp[ax] - r), 2) * 4, hi = child_cell(int(pp[ax] + 1 + r), 2) * 4;
                        am[ax] = (lo <= org[ax] && org[ax] <= hi ? 1u : 0u) |
                                 (lo <= org[ax] + 4 && org[ax] + 4 <= hi ? 2u : 0u);
                    }
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        if ((am[0] >> (k & 1)) & (am[1] >> ((k >> 1) & 1)) & (am[2] >> (k >> 2)) & 1u) um |= 1u << k;
                    S.pt[q] = make_float4(p.x, p.y, p.z, __uint_as_float(pos));
                }
                S.um[q] = uint8_t(um);
            }
            __syncthreads();
            // this wave's candidates, in candidate order
            uint32_t n = 0;
            const int ngrp = int((min(uint32_t(GB_CHUNK), wend - c0) + 63) / 64);
            for (int grp = 0; grp < ngrp; grp++) {
                const bool on = (S.um[grp * 64 + lane] >> wave) & 1u;
                const u64 bal = __ballot(on);
                if (on) S.lst[wave][n + uint32_t(__popcll(bal & lower))] = uint16_t(grp * 64 + lane);
                n += uint32_t(__popcll(bal));
            }
            __builtin_amdgcn_wave_barrier();
            for (uint32_t k = 0; k < n; k += 4) {
                const uint32_t j = k + uint32_t(g);
                if (j < n) {
                    const float4 pe = S.pt[S.lst[wave][j]];
                    const float px = pe.x, py = pe.y, pz = pe.z;
                    const uint32_t id = __float_as_uint(pe.w);
                    const float dy = py - fy, dz = pz - fz;
                    const float dyz = __builtin_fmaf(dy, dy, dz * dz);
#pragma unroll
                    for (int x = 0; x < 4; x++) {
                        const float dx = px - fx[x];
                        best_offer(B[x], __builtin_fmaf(dx, dx, dyz), px, py, pz, id, fx[x], fy, fz, a.sorted);
                    }
                }
            }
            __syncthreads();
        }
    }
    // merge the four candidate groups of each row
#pragma unroll
    for (int sh = 16; sh <= 32; sh <<= 1) {
#pragma unroll
        for (int x = 0; x < 4; x++) {
            const float obt = __shfl_xor(B[x].bt, sh, 64);
            const uint32_t oi = __shfl_xor(B[x].id, sh, 64);
            if (oi != 0xFFFFFFFFu) {
                const float4 q = a.sorted[oi];
                best_offer(B[x], __uint_as_float(__float_as_uint(obt) - NEAR_ULPS), q.x, q.y, q.z, oi, fx[x], fy, fz,
                           a.sorted);
            }
        }
    }
    if (g != 0) return;
    AT* atlas = reinterpret_cast<AT*>(a.atlas);
    const int vy = sy + ry, vz = sz + rz;
    const u64 aoff = atlas_offset(t, bn->mValue[0] + (sx - bn->mPos[0]), bn->mValue[1] + (vy - bn->mPos[1]),
                                  bn->mValue[2] + (vz - bn->mPos[2]));
#pragma unroll
    for (int x = 0; x < 4; x++) {
        float val = a.accumulate ? load_atlas<AT>(atlas, aoff + x) : (a.levelset ? 3.0f : 0.0f);
        if (B[x].id != 0xFFFFFFFFu) {
            const float4 q = a.sorted[B[x].id];
            const float f = falloff_accurate(q.x, q.y, q.z, sx + x, vy, vz, r);
            val = a.levelset ? fminf(val, -f) : fmaxf(val, f);
        }
        store_atlas<AT>(atlas, aoff + x, val);
    }
}
GVDB_MULTI_T(k_bucket_gather, typename AT, AT, BucketGatherArgs, GB_THREADS)

// GatherDensity by scattering: every candidate point offers itself to the
// voxels of the brick within the radius of it, in a copy of the brick held
// in LDS, and the brick is written once.  The result is
// max(0, max_p (r - |p - v|)) over the points p of the subcells around the
// brick, the value the reference's entry loop
// (kernels/cuda_gvdb_particles.cuh:724-728) computes, since a point with
// r - |p - v| > 0 is in the entry list of v's subcell.  Work per point is its
// box of (2r)^3 voxels instead of every voxel of the subcells it reaches.
//
// Exact squared distances in integers: all candidates of a brick lie at or
// above pmin = (brick origin - window) - 1 on every axis.  With
// pmin >= 1/32 they are multiples of 2^-s, s = 23 - floor(log2 pmin) <= 28,
// so X = (p - v) 2^s (p - v taken exactly as a TwoSum pair) is an exact
// integer, below 2^31 for |p - v| < 8; D = X^2 + Y^2 + Z^2 (u64, under
// 3 * 2^62) is the exact d^2 scaled by 2^2s.  Each voxel keeps the least D
// (ds_min_u64: order free, so deterministic; equal D are equal values), and
// r - sqrt(D) 2^-s is evaluated once per voxel.  Bricks whose window reaches
// below 1/32 take the float path: the accurate falloff of every pair, kept
// as a max through the same u64 slots.
//
// Per chunk of GB_CHUNK candidates: every thread loads two (independent
// loads in flight), the ones whose box meets the brick are compacted into
// LDS, then work items (candidate, z slice of its box) spread over the
// workgroup.  The float bounds of a box are supersets of the voxels within r
// (rounding is monotone and integers are representable).  LDS rows of a z
// slice are padded to 65 slots so the slices of one candidate fall in
// different banks.
static const int PG_ZSTRIDE = 65;
// Radii whose box has at most PG_ZMAX z slices (2r + 1/16 < PG_ZMAX: r <= 2)
// list only the slices inside the brick as work items; larger ones walk every
// slice of every candidate.
static const int PG_ZMAX = 5;
// NT threads per brick: PG_SMALL = 128 (two waves; measured faster than 64
// and 256 on the medium workload) when the (2 + 2R)^3 cells fit one per
// thread (R = 1: 64), else 256 (R = 2: 216 cells).
static const int PG_SMALL = 128;
// Listed variants: the cell of each of the first PG_OWN candidates comes from
// a table filled at set-up (one LDS byte read instead of a binary search over
// the 64 cell starts); later candidates search.
static const int PG_OWN = 512;
template <int NT, int LS>
struct PGShared {
    // the listed variant runs with R = 1 only: (2 + 2R)^3 = 64 cells
    uint32_t cell_off[LS ? 64 : NT];
    uint32_t cell_start[LS ? 64 : NT];
    int nbr[27];
    uint32_t wsum[NT / 64];
    uint32_t ncand, nitems;
    u64 key[8 * PG_ZSTRIDE];
    typename std::conditional<LS != 0, float3, float4>::type pt[2 * NT];
    uint16_t item[LS ? 2 * NT * PG_ZMAX : 2];   // candidate << 3 | z slice in the brick
    uint8_t owner[LS ? PG_OWN : 4];             // cell of candidate f < PG_OWN
};

// r - sqrt(D) 2^-s, accurate to a few ulps: D split into float hi + lo
// (exact to 2^-48), r^2 - d^2 in double-float, divided by r + d
__device__ inline float falloff_from_d2(u64 D, int s, float r) {
#pragma clang fp contract(off)
    const float dh = float(D);
    const long long rem = (long long)(D - u64(dh));
    const float hi = __builtin_ldexpf(dh, -2 * s);
    const float lo = __builtin_ldexpf(float(rem), -2 * s);
    const float r2h = r * r;
    const float r2l = __builtin_fmaf(r, r, -r2h);
    const float num = (r2h - hi) + (r2l - lo);
    // the hardware square root and reciprocal (1 ulp each): the quotient
    // keeps a relative error of a few ulps, as num does, down to the edge
    return num * __builtin_amdgcn_rcpf(r + __builtin_amdgcn_sqrtf(hi));
}

// LS = 0: every slice of every candidate, variable loops.  LS = 5: the slices
// inside the brick listed, rows and columns in fixed unrolled loops of 5
// (2r + 1/16 < 5).  LS = 4 (r <= 2): on bricks with exact integer distances
// each axis starts at the first integer above p - r (found exactly), so four
// slots cover every voxel with d < r; the voxels this leaves out have
// D >= r^2 2^2s, whose value r - d is <= 0 and stores 0 either way.
template <int NT, typename AT, int LS>
__device__ inline void k_push_gather_body(const BucketGatherArgs& a, uint32_t bid, uint32_t) {
    __shared__ PGShared<NT, LS> S;
    const Tree& t = a.t;
    const int brick = int(bid);
    const Node* bn = t.nodes[0] + brick;
    const int nb_mine = gb_nbr(a.nbr, brick);     // in flight during the checks
    if (bn->mParent == ID_UNDEF64) return;        // unlinked leaves have no subcells
    if (!bn->mFlags || bn->mValue[0] < 0) return;  // no atlas slot to write (uniform per brick)
    const int tid = threadIdx.x, lane = tid & 63;
    // voxel v = tid + 256 j (j = 0, 1) of the brick: x = v & 7, y = (v >> 3) & 7, z = v >> 6
    for (int v = tid; v < 512; v += NT) S.key[(v >> 6) * PG_ZSTRIDE + (v & 63)] = ~0ull;
    const uint32_t T = gb_setup<NT>(a.R, nb_mine, a.bcnt, a.boff, S);
    __syncthreads();                               // cell table and brick copy ready
    const float r = a.radius;
    const float r2cut = r * r * (1.0f + 0x1p-20f);
    const int span = 2 + 2 * a.R, ncell = span * span * span;
    if constexpr (LS != 0) {
        if (tid < ncell) {
            const uint32_t b = S.cell_start[tid], e = tid + 1 < ncell ? S.cell_start[tid + 1] : T;
            for (uint32_t f = b; f < min(e, uint32_t(PG_OWN)); f++) S.owner[f] = uint8_t(tid);
        }
        __syncthreads();
    }
    const int ox = bn->mPos[0], oy = bn->mPos[1], oz = bn->mPos[2];
    const float bx0 = float(ox), by0 = float(oy), bz0 = float(oz);
    const float bx1 = float(ox + 7), by1 = float(oy + 7), bz1 = float(oz + 7);
    // fixed-point scale of the brick (see above); 0 selects the float path
    const float pmin = float(min(ox, min(oy, oz)) - a.R * 4 - 1);
    int sh = 0;
    if (pmin >= 0x1p-5f) sh = 23 - (int((__float_as_uint(pmin) >> 23) & 0xFF) - 127);
    const bool exact = sh > 0;
    const float scale = __builtin_ldexpf(1.0f, sh);
    const int step = 1 << (exact ? sh : 0);
    const u64 dcut = exact ? u64(__builtin_ldexpf(r2cut, 2 * sh)) : 0ull;
    // integers in [fl(p - r), fl(p + r)]: the interval exceeds 2r by two
    // roundings at most (under 1/32 each for coordinates below 2^18)
    const uint32_t nz = uint32_t(2.0f * r + 0.0625f) + 1u;
    const uint32_t nzinv = uint32_t(0xFFFFFFFFu / nz) + 1u;
    constexpr bool listed = LS != 0;   // the host picks LS from the radius
    const bool tight = LS == 4 && exact;
    const u64 lower = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    for (uint32_t c0 = 0; c0 < T; c0 += (2 * NT)) {
        float4 p[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t f = c0 + uint32_t(h * NT + tid);
            if (f < T) {
                int lo = 0;
                if (LS != 0 && f < uint32_t(PG_OWN)) {
                    lo = S.owner[f];
                } else {
                    int hi = ncell - 1;            // last cell starting at or before f
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (S.cell_start[mid] <= f) lo = mid;
                        else hi = mid - 1;
                    }
                }
                p[h] = a.sorted[S.cell_off[lo] + (f - S.cell_start[lo])];
            }
        }
        if (tid == 0) S.ncand = S.nitems = 0;
        __syncthreads();                           // counters cleared; previous chunk done with pt
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t f = c0 + uint32_t(h * NT + tid);
            bool keep = false;
            int zlo = 0, nzv = 0;
            if (f < T) {
                const float4 q = p[h];
                keep = q.x - r <= bx1 && q.x + r >= bx0 && q.y - r <= by1 && q.y + r >= by0 &&
                       q.z - r <= bz1 && q.z + r >= bz0;
                if (keep && listed) {
                    // the slices z in [fl(q.z - r), fl(q.z + r)] inside the brick, at
                    // most nz of them from the first (so nzv <= nz <= PG_ZMAX); tight:
                    // from the first integer above q.z - r, at most four
                    const int zc = tight ? first_above(q.z, r) : int(ceilf(q.z - r));
                    const int cap = tight ? 4 : int(min(nz, uint32_t(PG_ZMAX)));
                    zlo = max(zc, oz);
                    nzv = max(min(min(int(floorf(q.z + r)), zc + cap - 1), oz + 7) - zlo + 1, 0);
                    keep = nzv > 0;
                }
            }
            const u64 bal = __ballot(keep);
            uint32_t base = 0;
            if constexpr (LS != 0) {
                // lanes' slice counts (0..5) summed below each lane, from the bits
                const u64 b0 = __ballot(nzv & 1), b1 = __ballot(nzv & 2), b2 = __ballot(nzv & 4);
                const uint32_t ipre = uint32_t(__popcll(b0 & lower) + 2 * __popcll(b1 & lower) + 4 * __popcll(b2 & lower));
                const uint32_t itot = uint32_t(__popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2));
                uint32_t ibase = 0;
                if (lane == 0 && bal) {
                    base = atomicAdd(&S.ncand, uint32_t(__popcll(bal)));
                    ibase = atomicAdd(&S.nitems, itot);
                }
                base = __shfl(base, 0, 64);
                ibase = __shfl(ibase, 0, 64);
                if (keep) {
                    const uint32_t ci = base + uint32_t(__popcll(bal & lower));
                    S.pt[ci] = make_float3(p[h].x, p[h].y, p[h].z);
                    uint16_t* dst = S.item + ibase + ipre;
                    const uint32_t w0 = ci << 3 | uint32_t(zlo - oz);
#pragma nounroll
                    for (int k = 0; k < nzv; k++) dst[k] = uint16_t(w0 + uint32_t(k));
                }
            } else {
                if (lane == 0 && bal) base = atomicAdd(&S.ncand, uint32_t(__popcll(bal)));
                base = __shfl(base, 0, 64);
                if (keep) S.pt[base + uint32_t(__popcll(bal & lower))] = p[h];
            }
        }
        __syncthreads();
        const uint32_t items = listed ? S.nitems : S.ncand * nz;
        for (uint32_t it = uint32_t(tid); it < items; it += NT) {
            uint32_t c;
            int z;
            float4 q;
            if constexpr (LS != 0) {
                const uint32_t w = S.item[it];
                c = w >> 3;
                z = oz + int(w & 7u);
                const auto q3 = S.pt[c];
                q = make_float4(q3.x, q3.y, q3.z, 0.0f);
            } else {
                c = __umulhi(it, nzinv);
                if (c * nz > it) c--;
                const int zi = int(it - c * nz);
                q = S.pt[c];
                z = int(ceilf(q.z - r)) + zi;
                if (z < oz || z > oz + 7 || float(z) > q.z + r) continue;
            }
            const int y0 = max(tight ? first_above(q.y, r) : int(ceilf(q.y - r)), oy), y1 = min(int(floorf(q.y + r)), oy + 7);
            const int x0 = max(tight ? first_above(q.x, r) : int(ceilf(q.x - r)), ox), x1 = min(int(floorf(q.x + r)), ox + 7);
            if (y0 > y1 || x0 > x1) continue;
            u64* kz = S.key + (z - oz) * PG_ZSTRIDE - ox;
            const float dz = q.z - float(z);
            if (exact) {
                // p - v = hi + lo exactly (TwoSum); both are multiples of 2^-s
                float ez, ey, ex;
                const float hz = two_diff(q.z, float(z), ez), hy = two_diff(q.y, float(y0), ey),
                            hx = two_diff(q.x, float(x0), ex);
                const int Z = int(hz * scale) + int(ez * scale);
                const int X0 = int(hx * scale) + int(ex * scale);
                int Y = int(hy * scale) + int(ey * scale);
                const long long Z2 = (long long)Z * Z;
                if constexpr (LS != 0) {
                    // at most LS rows and columns (the box bound of the slices
                    // holds per axis for coordinates below 2^18): fixed, unrolled
                    u64* k0 = kz + (y0 - oy) * 8 + x0;
#pragma unroll
                    for (int j = 0; j < LS; j++) {
                        const int Yj = Y - j * step;
                        const u64 YZ = u64((long long)Yj * Yj + Z2);
                        if (y0 + j > y1 || YZ >= dcut) continue;
#pragma unroll
                        for (int i = 0; i < LS; i++) {
                            const int Xi = X0 - i * step;
                            const u64 D = u64((long long)Xi * Xi) + YZ;
                            if (x0 + i <= x1 && D < dcut) atomicMin(k0 + j * 8 + i, D);
                        }
                    }
                } else {
                    for (int y = y0; y <= y1; y++, Y -= step) {
                        const u64 YZ = u64((long long)Y * Y + Z2);
                        if (YZ >= dcut) continue;
                        u64* kr = kz + (y - oy) * 8;
                        int X = X0;
                        for (int x = x0; x <= x1; x++, X -= step) {
                            const u64 D = u64((long long)X * X) + YZ;
                            if (D < dcut) atomicMin(kr + x, D);
                        }
                    }
                }
            } else {
                const float dz2 = dz * dz;
                for (int y = y0; y <= y1; y++) {
                    const float dy = q.y - float(y);
                    const float dyz = __builtin_fmaf(dy, dy, dz2);
                    if (!(dyz < r2cut)) continue;
                    u64* kr = kz + (y - oy) * 8;
                    for (int x = x0; x <= x1; x++) {
                        const float dx = q.x - float(x);
                        if (__builtin_fmaf(dx, dx, dyz) < r2cut) {
                            const float v = falloff_inl(q.x, q.y, q.z, x, y, z, r);
                            if (v > 0.0f) atomicMin(kr + x, u64(0x7F800000u - __float_as_uint(v)));
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();
    // the brick's atlas slot found once (its divisions by the padded side are
    // per brick, not per voxel); a voxel adds its padded-local index
    AT* atlas = reinterpret_cast<AT*>(a.atlas) + slot_of(t, bn->mValue) * u64(t.brick_vox);
    const int pbr = t.brick_res, ap = t.apron;
    AT* F = a.faces ? reinterpret_cast<AT*>(a.faces) + u64(brick) * FACE_VALS : nullptr;
    for (int v = tid; v < 512; v += NT) {
        const int lx = v & 7, ly = (v >> 3) & 7, lz = v >> 6;
        const u64 k = S.key[lz * PG_ZSTRIDE + (v & 63)];
        float val = 0.0f;
        if (k != ~0ull)
            val = exact ? fmaxf(0.0f, falloff_from_d2(k, sh, r)) : __uint_as_float(0x7F800000u - uint32_t(k));
        store_atlas<AT>(atlas, u64(((lz + ap) * pbr + ly + ap) * pbr + lx + ap), val);
        if (F) {
            // the brick's boundary layers, compact, for the apron pass
            if (lx == 0) store_atlas<AT>(F, 0 * 64 + lz * 8 + ly, val);
            if (lx == 7) store_atlas<AT>(F, 1 * 64 + lz * 8 + ly, val);
            if (ly == 0) store_atlas<AT>(F, 2 * 64 + lz * 8 + lx, val);
            if (ly == 7) store_atlas<AT>(F, 3 * 64 + lz * 8 + lx, val);
            if (lz == 0) store_atlas<AT>(F, 4 * 64 + ly * 8 + lx, val);
            if (lz == 7) store_atlas<AT>(F, 5 * 64 + ly * 8 + lx, val);
        }
    }
}

// Written out (not GVDB_MULTI_T) to ask for eight waves per SIMD (at most 64
// VGPRs).
// MULTI = false: one grid, its record by value (launched when the launch
// covers one grid), so the single-grid kernel does not carry the registers of
// the multi-grid record lookup (63 VGPRs without spills instead of 64 with).
template <int NT, typename AT, int LS, bool MULTI>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_push_gather(BucketGatherArgs a0,
                                                                                              MultiMap m) {
    const uint32_t bi = m.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    if (!MULTI) {
        k_push_gather_body<NT, AT, LS>(a0, bi, gridDim.x);
        return;
    }
    // several grids: first[] and the records in device memory (set up with
    // inline_ok = false); the record copied once, from a wave-uniform index
    const int g = __builtin_amdgcn_readfirstlane(multi_grid_dev(m, bi));
    const BucketGatherArgs a = multi_rec_dev<BucketGatherArgs>(m, g);
    const uint32_t f0 = m.first[g], f1 = m.first[g + 1];
    k_push_gather_body<NT, AT, LS>(a, bi - f0, f1 - f0);
}

template <int NT, typename AT, int LS>
static void launch_push_gather(dim3 grid, hipStream_t s, const BucketGatherArgs& a0, const MultiMap& m) {
    if (m.G > 1) hipLaunchKernelGGL((k_push_gather<NT, AT, LS, true>), grid, dim3(NT), 0, s, a0, m);
    else hipLaunchKernelGGL((k_push_gather<NT, AT, LS, false>), grid, dim3(NT), 0, s, a0, m);
}

// The push gather's z slices per candidate, as the kernel computes it.
static inline uint32_t pg_slices(float r) { return uint32_t(2.0f * r + 0.0625f) + 1u; }

hipError_t launch_bucket_gather_multi(const BucketGatherArgs* a, int G, Arena* ar, hipStream_t s) {
    // variants: scatter kernel (no accumulate) with 128 threads and 4, 5 or any
    // voxels per axis, or 256 threads, or the nearest-entry kernel (accumulate:
    // there the reference's max also takes the negative r - |p - v| of entries
    // beyond the radius); FP32 or FP16
    for (int variant = 0; variant < 10; variant++) {
        const bool half = variant & 1;
        // 0: scatter 128, LS 4; 1: scatter 128, LS 5; 2: scatter 128; 3: scatter 256; 4: nearest entry
        const int kind = variant >> 1;
        std::vector<BucketGatherArgs> sel;
        std::vector<uint32_t> nb;
        for (int g = 0; g < G; g++) {
            if (a[g].t.res[0] != 8 || a[g].R < 1 || a[g].R > 2) return hipErrorInvalidValue;
            // listed slices: at most PG_ZMAX per axis, coordinates below 2^18
            const bool listed = pg_slices(a[g].radius) <= uint32_t(PG_ZMAX) &&
                                a[g].t.range[a[g].t.levels - 1] <= (1 << 18);
            const int k = (a[g].accumulate || a[g].levelset) ? 4
                          : a[g].R == 2                      ? 3
                          : !listed                          ? 2
                          : a[g].radius <= 2.0f              ? 0
                                                             : 1;
            if ((a[g].atlas_type == T_HALF) != half || k != kind || a[g].nleaves <= 0) continue;
            sel.push_back(a[g]);
            nb.push_back(uint32_t(a[g].nleaves));
        }
        if (sel.empty()) continue;
        MultiLaunch L;
        // the push gather's multi-grid kernels read device tables only
        hipError_t e = multi_setup(ar, sel.data(), nb.data(), int(sel.size()), s, L, kind == 4);
        if (e != hipSuccess) return e;
        if (L.blocks == 0) continue;
        const dim3 grid(L.blocks);
        if (kind < 4) L.map.xcd = knob(KNOB_XCD_ORDER);
        // the listed-slice variants run with the tuned threads per brick
        const bool wide = knob(KNOB_GATHER_THREADS) == 256;
        switch (variant) {
        case 0:
            if (wide) launch_push_gather<256, float, 4>(grid, s, sel[0], L.map);
            else launch_push_gather<PG_SMALL, float, 4>(grid, s, sel[0], L.map);
            break;
        case 1:
            if (wide) launch_push_gather<256, __half, 4>(grid, s, sel[0], L.map);
            else launch_push_gather<PG_SMALL, __half, 4>(grid, s, sel[0], L.map);
            break;
        case 2:
            if (wide) launch_push_gather<256, float, 5>(grid, s, sel[0], L.map);
            else launch_push_gather<PG_SMALL, float, 5>(grid, s, sel[0], L.map);
            break;
        case 3:
            if (wide) launch_push_gather<256, __half, 5>(grid, s, sel[0], L.map);
            else launch_push_gather<PG_SMALL, __half, 5>(grid, s, sel[0], L.map);
            break;
        case 4: launch_push_gather<PG_SMALL, float, 0>(grid, s, sel[0], L.map); break;
        case 5: launch_push_gather<PG_SMALL, __half, 0>(grid, s, sel[0], L.map); break;
        case 6: launch_push_gather<256, float, 0>(grid, s, sel[0], L.map); break;
        case 7: launch_push_gather<256, __half, 0>(grid, s, sel[0], L.map); break;
        case 8: hipLaunchKernelGGL(k_bucket_gather<float>, grid, dim3(GB_THREADS), 0, s, sel[0], L.map); break;
        default: hipLaunchKernelGGL(k_bucket_gather<__half>, grid, dim3(GB_THREADS), 0, s, sel[0], L.map); break;
        }
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ------------------------------------------------------- colour gather
// gvdbGatherDensity's colour (kernels/cuda_gvdb_particles.cuh:717-731; the
// level set's :757-771 is the same): starting from (0, 0, 0, 1), or from the
// voxel's colour with bAccumulate, it adds every entry's colour / 255 over the
// subcell's list and divides by the entry count (+1 with bAccumulate), then
// stores x * 255 truncated to uchar (clamped: an empty list divides by zero,
// storing 0 for NaN and 255 for the alpha's +inf).  Here the sums are exact
// integers, so the stored byte is (255 [alpha only] + sum) / count, or
// (old + sum) / (count + 1), in integer division: the mean of the list's
// colours, independent of the order the entries were inserted in (the
// reference adds floats in atomic-insertion order).  One workgroup per brick
// reads the candidates of the brick's (2 + 2R)^3 cells, as the density
// gather does, and counts a candidate for each of the brick's subcells whose
// list the insert rule puts it in: int(w - r)/sc*sc <= s <= int(w + 1 + r)/sc*sc
// on every axis (kernels/cuda_gvdb_particles.cuh:237-242).
static const int CG_THREADS = 256;
struct CGShared {
    uint32_t cell_off[CG_THREADS];
    uint32_t cell_start[CG_THREADS];
    int nbr[27];
    uint32_t wsum[CG_THREADS / 64];
    uint32_t sum[8][5];   // r, g, b, a, entries per subcell of the brick
};

__device__ inline void k_gather_color_body(const ColorArgs& a, uint32_t bid, uint32_t) {
    __shared__ CGShared S;
    const Tree& t = a.t;
    const int brick = int(bid);
    const Node* bn = t.nodes[0] + brick;
    const int nb_mine = gb_nbr(a.nbr, brick);
    if (bn->mParent == ID_UNDEF64) return;        // unlinked leaves have no subcells
    if (!bn->mFlags || bn->mValue[0] < 0) return;  // no atlas slot to write (uniform per brick)
    const int tid = threadIdx.x;
    if (tid < 40) (&S.sum[0][0])[tid] = 0;
    const uint32_t T = gb_setup<CG_THREADS>(a.R, nb_mine, a.bcnt, a.boff, S);
    __syncthreads();
    const int span = 2 + 2 * a.R, ncell = span * span * span;
    const int sc = a.scs, sd = t.brickwid / sc;   // subcells per axis (2 for 8^3 bricks, sc 4)
    const float r = a.radius;
    for (uint32_t f = uint32_t(tid); f < T; f += CG_THREADS) {
        int lo = 0, hi = ncell - 1;                // last cell starting at or before f
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (S.cell_start[mid] <= f) lo = mid;
            else hi = mid - 1;
        }
        const float4 q = a.sorted[S.cell_off[lo] + (f - S.cell_start[lo])];
        const float w[3] = {q.x, q.y, q.z};
        uint32_t m[3];                              // subcells of the brick along each axis whose list takes it
        for (int ax = 0; ax < 3; ax++) {
            const int s0 = int(w[ax] - r) / sc * sc, s1 = int(w[ax] + 1.0f + r) / sc * sc;
            m[ax] = 0;
            for (int k = 0; k < sd; k++) {
                const int st = bn->mPos[ax] + k * sc;
                if (s0 <= st && st <= s1) m[ax] |= 1u << k;
            }
        }
        if (!(m[0] && m[1] && m[2])) continue;
        const uint32_t c = __float_as_uint(q.w);
        for (int kz = 0; kz < sd; kz++)
            for (int ky = 0; ky < sd; ky++)
                for (int kx = 0; kx < sd; kx++) {
                    if (!((m[0] >> kx) & (m[1] >> ky) & (m[2] >> kz) & 1u)) continue;
                    uint32_t* sm = S.sum[(kz * sd + ky) * sd + kx];
                    atomicAdd(sm + 0, c & 0xFFu);
                    atomicAdd(sm + 1, (c >> 8) & 0xFFu);
                    atomicAdd(sm + 2, (c >> 16) & 0xFFu);
                    atomicAdd(sm + 3, c >> 24);
                    atomicAdd(sm + 4, 1u);
                }
    }
    __syncthreads();
    uint32_t* atlas = reinterpret_cast<uint32_t*>(a.atlas) + slot_of(t, bn->mValue) * u64(t.brick_vox);
    const int w = t.brickwid, pbr = t.brick_res, ap = t.apron;
    for (int v = tid; v < w * w * w; v += CG_THREADS) {
        const int lx = v % w, ly = (v / w) % w, lz = v / (w * w);
        const uint32_t* sm = S.sum[((lz / sc) * sd + ly / sc) * sd + lx / sc];
        const u64 off = u64(((lz + ap) * pbr + ly + ap) * pbr + lx + ap);
        const uint32_t n = sm[4];
        uint32_t out = 0;
        if (a.accumulate) {
            const uint32_t old = atlas[off];
            for (int ch = 0; ch < 4; ch++) out |= (((old >> (8 * ch)) & 0xFFu) + sm[ch]) / (n + 1u) << (8 * ch);
        } else if (n == 0) {
            out = 0xFF000000u;
        } else {
            for (int ch = 0; ch < 4; ch++) out |= min((sm[ch] + (ch == 3 ? 255u : 0u)) / n, 255u) << (8 * ch);
        }
        atlas[off] = out;
    }
}

GVDB_MULTI(k_gather_color, ColorArgs, CG_THREADS)

hipError_t launch_gather_color_multi(const ColorArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> nb(G);
    for (int g = 0; g < G; g++) {
        if (a[g].t.res[0] != 8 || a[g].R < 1 || a[g].R > 2 || a[g].scs <= 0 || a[g].t.brickwid % a[g].scs ||
            a[g].t.brickwid / a[g].scs > 2 || (2 + 2 * a[g].R) * (2 + 2 * a[g].R) * (2 + 2 * a[g].R) > CG_THREADS)
            return hipErrorInvalidValue;
        nb[g] = uint32_t(a[g].nleaves > 0 ? a[g].nleaves : 0);
    }
    MultiLaunch L;
    hipError_t e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    hipLaunchKernelGGL(k_gather_color, dim3(L.blocks), dim3(CG_THREADS), 0, s, a[0], L.map);
    return hipGetLastError();
}

GVDB_MULTI(k_bucket_points, BucketArgs, 256)
GVDB_MULTI(k_bucket_scatter, ScatterArgs, 256)
GVDB_MULTI_T(k_pull, bool WRITE, WRITE, PullArgs, PULL_THREADS)

hipError_t launch_bucket_points_multi(const BucketArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> items(G);
    for (int g = 0; g < G; g++) items[g] = uint32_t(a[g].pts.n > 0 ? a[g].pts.n : 0);
    return launch_per_item<BucketArgs, k_bucket_points>(a, items.data(), G, ar, s, 256);
}

hipError_t launch_bucket_scatter_multi(const ScatterArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> items(G);
    for (int g = 0; g < G; g++) items[g] = uint32_t(a[g].pts.n > 0 ? a[g].pts.n : 0);
    return launch_per_item<ScatterArgs, k_bucket_scatter>(a, items.data(), G, ar, s, 256);
}

int pull_cells(int sd, int R) { int span = sd + 2 * R; return span * span * span; }
int pull_max_cells() { return PULL_MAX_CELLS; }
int pull_max_spb() { return PULL_MAX_SPB; }

hipError_t launch_pull_multi(int write, const PullArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> nb(G);
    for (int g = 0; g < G; g++) nb[g] = uint32_t(a[g].nleaves > 0 ? a[g].nleaves : 0);
    MultiLaunch L;
    hipError_t e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    if (write) hipLaunchKernelGGL(k_pull<true>, dim3(L.blocks), dim3(PULL_THREADS), 0, s, a[0], L.map);
    else hipLaunchKernelGGL(k_pull<false>, dim3(L.blocks), dim3(PULL_THREADS), 0, s, a[0], L.map);
    return hipGetLastError();
}

}  // namespace gvdb


