// keysort.hip -- the full topology build over sorted brick keys, for point
// sets whose level-0 brick box is too large for bitmaps (ActivateBricksGPU /
// ActivateExtraBricksGPU, src/gvdb_volume_gvdb.cpp:1181-1331 and :935-1018:
// the reference's own form, keys of every point's brick and of its dilation
// box, sorted and made unique per level).  A key is the x-major index of a
// cell in the level's brick box (the box of the bitmap build), so sorted keys
// give the bitmap build's node order; both builds give the same pools.
//
// The sort is a least-significant-digit radix sort over 8-bit digits: per
// pass, a digit histogram per 2048-key tile (LDS atomics), one exclusive scan
// over the digit-major table, and a stable scatter in which each wave ranks
// its keys among equal digits with ballots (eight ballots give the lanes
// sharing a lane's digit).
#include "kernels.h"
#include "tree_device.h"
#include <algorithm>

namespace gvdb {

static const int RS_THREADS = 256;
static const int RS_ROUNDS = 8;
static const int RS_TILE = RS_THREADS * RS_ROUNDS;

static uint32_t grid_of(u64 n, u64 per, u64 cap = 1u << 20) {
    u64 b = (n + per - 1) / per;
    return uint32_t(std::max<u64>(1, std::min<u64>(b, cap)));
}

__device__ inline bool in_box(const BitDomain& d, int x, int y, int z, u64& idx) {
    const int a = x - d.lo[0], b = y - d.lo[1], c = z - d.lo[2];
    if (a < 0 || b < 0 || c < 0 || a >= d.dim[0] || b >= d.dim[1] || c >= d.dim[2]) return false;
    idx = (u64(a) * d.dim[1] + b) * d.dim[2] + c;
    return true;
}

__device__ inline void box_coords(const BitDomain& d, u64 idx, int& x, int& y, int& z) {
    const u64 yz = u64(d.dim[1]) * d.dim[2];
    const u64 a = idx / yz, r = idx - a * yz, b = r / d.dim[2];
    x = int(a) + d.lo[0];
    y = int(b) + d.lo[1];
    z = int(r - b * d.dim[2]) + d.lo[2];
}

// Per point: its own brick's key (or the sentinel) and the number of bricks
// of its dilation box inside the level-0 box.  The brick tests are the bitmap
// build's (k_mark_points): the own brick floor(w / range0), the dilation
// box [int(int(w) - r), int(int(w) + 2r)] / range0 with C truncation.
__device__ inline bool key_point(const KeyGenArgs& a, int i, float w[3]) {
    const float* p = reinterpret_cast<const float*>(a.pts.base + u64(i) * a.pts.stride + a.pts.off);
    w[0] = p[0] + a.o[0];
    w[1] = p[1] + a.o[1];
    w[2] = p[2] + a.o[2];
    if (!isfinite(w[0]) || !isfinite(w[1]) || !isfinite(w[2])) return false;
    return !(w[0] < 0 || w[1] < 0 || w[2] < 0);
}

__device__ inline void dil_box(const KeyGenArgs& a, const float w[3], int lo[3], int hi[3]) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int iw = int(w[k]);
        lo[k] = int(float(iw) - a.radius) / a.range0;
        hi[k] = int(float(iw) + 2.0f * a.radius) / a.range0;
    }
}

__global__ __launch_bounds__(256) void k_key_count(KeyGenArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.pts.n) return;
    float w[3];
    u64 own = a.sentinel;
    uint32_t cnt = 0;
    if (key_point(a, i, w)) {
        int c[3];
#pragma unroll
        for (int k = 0; k < 3; k++) c[k] = int(floorf(w[k] * a.inv_range0));
        u64 idx;
        if (in_box(a.d0, c[0], c[1], c[2], idx)) own = idx;
        if (a.dilate) {
            int lo[3], hi[3];
            dil_box(a, w, lo, hi);
            // the box clipped to the level-0 box, as the per-brick test does
            u64 n = 1;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int l = max(lo[k], a.d0.lo[k]), h = min(hi[k], a.d0.lo[k] + a.d0.dim[k] - 1);
                n *= h >= l ? u64(h - l + 1) : 0ull;
            }
            cnt = uint32_t(min(n, u64(0xFFFFFFFFu)));
        }
    }
    a.keysA[i] = own;
    a.cntB[i] = cnt;
}

__global__ __launch_bounds__(256) void k_key_emit(KeyGenArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.pts.n || a.cntB[i] == 0) return;
    float w[3];
    if (!key_point(a, i, w)) return;
    int lo[3], hi[3];
    dil_box(a, w, lo, hi);
    u64* out = a.keysB + a.offB[i];
    for (int x = max(lo[0], a.d0.lo[0]); x <= min(hi[0], a.d0.lo[0] + a.d0.dim[0] - 1); x++)
        for (int y = max(lo[1], a.d0.lo[1]); y <= min(hi[1], a.d0.lo[1] + a.d0.dim[1] - 1); y++)
            for (int z = max(lo[2], a.d0.lo[2]); z <= min(hi[2], a.d0.lo[2] + a.d0.dim[2] - 1); z++) {
                u64 idx;
                in_box(a.d0, x, y, z, idx);
                *out++ = idx;
            }
}

hipError_t launch_key_gen(const KeyGenArgs& a, bool emit, hipStream_t s) {
    if (a.pts.n <= 0) return hipSuccess;
    if (emit) hipLaunchKernelGGL(k_key_emit, dim3(grid_of(u64(a.pts.n), 256, 1u << 31)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_key_count, dim3(grid_of(u64(a.pts.n), 256, 1u << 31)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------ radix sort --
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const u64* keys, u64 n, int shift, uint32_t* hist,
                                                        uint32_t nblk) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const u64 base = u64(blockIdx.x) * RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const u64 i = base + u64(r) * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[uint32_t(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[u64(threadIdx.x) * nblk + blockIdx.x] = h[threadIdx.x];
}

// Keys of a tile in rounds of 256 consecutive keys; within a round a key's
// place among equal digits is (equal digits of lower waves) + (equal digits
// of lower lanes of its wave), after the tile's earlier rounds (run[]), after
// the digit's keys of lower tiles and lower digits (offs[], digit-major).
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const u64* in, u64* out, u64 n, int shift,
                                                           const uint32_t* offs, uint32_t nblk) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[RS_THREADS / 64][256];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    run[t] = offs[u64(t) * nblk + blockIdx.x];
    const u64 base = u64(blockIdx.x) * RS_TILE;
    const u64 below = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int r = 0; r < RS_ROUNDS; r++) {
#pragma unroll
        for (int w = 0; w < RS_THREADS / 64; w++) wc[w][t] = 0;
        __syncthreads();
        const u64 i = base + u64(r) * RS_THREADS + t;
        const bool ok = i < n;
        const u64 k = ok ? in[i] : 0ull;
        const uint32_t d = uint32_t(k >> shift) & 255u;
        u64 peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const u64 bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t lr = uint32_t(__popcll(peers & below));
        if (ok && lr == 0) wc[wave][d] = uint32_t(__popcll(peers));
        __syncthreads();
        if (ok) {
            uint32_t pos = run[d] + lr;
            for (int w = 0; w < wave; w++) pos += wc[w][d];
            out[pos] = k;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < RS_THREADS / 64; w++) add += wc[w][t];
        run[t] += add;
    }
}

u64 radix_hist_words(u64 n) { return 256 * u64(grid_of(n, RS_TILE, ~0ull)); }

hipError_t radix_sort_u64(u64* keys, u64* tmp, u64 n, int bits, uint32_t* hist, uint32_t* scan_tmp, hipStream_t s,
                          u64** result) {
    *result = keys;
    if (n <= 1) return hipSuccess;
    const uint32_t nblk = grid_of(n, RS_TILE, ~0ull);
    u64 *src = keys, *dst = tmp;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(RS_THREADS), 0, s, src, n, shift, hist, nblk);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = launch_scan_u32(hist, nullptr, 0, hist, u64(nblk) * 256, nullptr, scan_tmp, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_rs_scatter, dim3(nblk), dim3(RS_THREADS), 0, s, src, dst, n, shift, hist, nblk);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        std::swap(src, dst);
    }
    *result = src;
    return hipSuccess;
}

// ------------------------------------------------- unique, difference ----
__global__ __launch_bounds__(256) void k_flag_unique(const u64* k, u64 n, u64 sentinel, uint32_t* flags) {
    const u64 i = u64(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const u64 v = k[i];
    flags[i] = (v != sentinel && (i == 0 || k[i - 1] != v)) ? 1u : 0u;
}

__device__ inline bool sorted_find(const u64* a, uint32_t n, u64 v, uint32_t& at) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    at = lo;
    return lo < n && a[lo] == v;
}

__global__ __launch_bounds__(256) void k_flag_absent(const u64* b, u64 nb, const u64* a, uint32_t na, uint32_t* flags) {
    const u64 i = u64(blockIdx.x) * 256 + threadIdx.x;
    if (i >= nb) return;
    uint32_t at;
    flags[i] = sorted_find(a, na, b[i], at) ? 0u : 1u;
}

__global__ __launch_bounds__(256) void k_compact(const u64* k, u64 n, const uint32_t* flags, const uint32_t* pos,
                                                 u64* out) {
    const u64 i = u64(blockIdx.x) * 256 + threadIdx.x;
    if (i < n && flags[i]) out[pos[i]] = k[i];
}

hipError_t launch_flag_unique(const u64* k, u64 n, u64 sentinel, uint32_t* flags, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_flag_unique, dim3(grid_of(n, 256, ~0ull)), dim3(256), 0, s, k, n, sentinel, flags);
    return hipGetLastError();
}
hipError_t launch_flag_absent(const u64* b, u64 nb, const u64* a, uint32_t na, uint32_t* flags, hipStream_t s) {
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_flag_absent, dim3(grid_of(nb, 256, ~0ull)), dim3(256), 0, s, b, nb, a, na, flags);
    return hipGetLastError();
}
hipError_t launch_compact(const u64* k, u64 n, const uint32_t* flags, const uint32_t* pos, u64* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact, dim3(grid_of(n, 256, ~0ull)), dim3(256), 0, s, k, n, flags, pos, out);
    return hipGetLastError();
}

// Parent keys: cell >> logdim of the parent level, in the parent's box.
__global__ __launch_bounds__(256) void k_parent_keys(const u64* c, u64 n, BitDomain dc, int ld, BitDomain dp,
                                                     u64* out) {
    const u64 i = u64(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    int x, y, z;
    box_coords(dc, c[i], x, y, z);
    u64 idx = 0;
    in_box(dp, x >> ld, y >> ld, z >> ld, idx);   // parents of cells in the box lie in the parent's box
    out[i] = idx;
}

hipError_t launch_parent_keys(const u64* c, u64 n, const BitDomain& dc, int ld, const BitDomain& dp, u64* out,
                              hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_parent_keys, dim3(grid_of(n, 256, ~0ull)), dim3(256), 0, s, c, n, dc, ld, dp, out);
    return hipGetLastError();
}

// ------------------------------------------------------ nodes and links --
__device__ inline void list_node(Node* n, int lev, int x, int y, int z, int range, u64 clist) {
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

// Node i of a level: the i-th key of the level's own-brick list, then of its
// dilation-only list (the bitmap build's rank order).
__global__ __launch_bounds__(256) void k_emit_list(KeyLevel L, Node* nodes, int has_clist) {
    const u64 i = u64(blockIdx.x) * 256 + threadIdx.x;
    if (i >= u64(L.na) + L.ne) return;
    const u64 key = i < L.na ? L.A[i] : L.E[i - L.na];
    int x, y, z;
    box_coords(L.dom, key, x, y, z);
    list_node(nodes + i, L.lev, x, y, z, L.range, has_clist ? Elem(1, L.lev, i) : ID_UNDEFL);
}

hipError_t launch_emit_list(const KeyLevel& L, Node* nodes, int has_clist, hipStream_t s) {
    const u64 n = u64(L.na) + L.ne;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_emit_list, dim3(grid_of(n, 256, ~0ull)), dim3(256), 0, s, L, nodes, has_clist);
    return hipGetLastError();
}

// gvdbLinkBricks (kernels/cuda_gvdb_nodes.cuh:470-505): the parent's index is
// the place of the parent's key in the parent level's lists; the child-list
// slot is (z * res + y) * res + x of the child inside its parent.
__global__ __launch_bounds__(256) void k_link_list(KeyLevel C, KeyLevel P, int parent_is_root, int ld, Node* cnodes,
                                                   u64* pclists, u64 clist_len) {
    const u64 i = u64(blockIdx.x) * 256 + threadIdx.x;
    if (i >= u64(C.na) + C.ne) return;
    const u64 key = i < C.na ? C.A[i] : C.E[i - C.na];
    int c[3];
    box_coords(C.dom, key, c[0], c[1], c[2]);
    const int pc[3] = {c[0] >> ld, c[1] >> ld, c[2] >> ld};
    uint32_t pidx = 0;
    if (!parent_is_root) {
        u64 pk;
        if (!in_box(P.dom, pc[0], pc[1], pc[2], pk)) return;
        uint32_t at;
        if (sorted_find(P.A, P.na, pk, at)) pidx = at;
        else if (sorted_find(P.E, P.ne, pk, at)) pidx = P.na + at;
        else return;   // not reachable: every node's parent cell is marked
    }
    const int res = 1 << ld;
    const int l[3] = {c[0] - (pc[0] << ld), c[1] - (pc[1] << ld), c[2] - (pc[2] << ld)};
    const int bit = (l[2] * res + l[1]) * res + l[0];
    cnodes[i].mParent = Elem(0, C.lev + 1, pidx);
    pclists[u64(pidx) * clist_len + bit] = Elem(0, C.lev, i);
}

hipError_t launch_link_list(const KeyLevel& C, const KeyLevel& P, int parent_is_root, int ld, Node* cnodes,
                            u64* pclists, u64 clist_len, hipStream_t s) {
    const u64 n = u64(C.na) + C.ne;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_link_list, dim3(grid_of(n, 256, ~0ull)), dim3(256), 0, s, C, P, parent_is_root, ld, cnodes,
                       pclists, clist_len);
    return hipGetLastError();
}

// ------------------------------------------------- incremental build ----
// gvdbCalcIncreExtraBrickId (kernels/cuda_gvdb_particles.cuh:420-465) with
// keys instead of bitmap bits: the cells of every level in the point's box
// that the walk from the top node does not find, as (level << 56 | cell key);
// the count pass also sets the markers of the leaves it finds (the bitmap
// form's k_incre_mark).
__global__ __launch_bounds__(256) void k_incre_keys(IncreMarkArgs a, uint32_t* cnt, const uint32_t* off, u64* keys,
                                                    int emit) {
    const Tree& t = a.t;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.pts.n) return;
    const float* p = reinterpret_cast<const float*>(a.pts.base + u64(i) * a.pts.stride + a.pts.off);
    const float w[3] = {p[0] + a.o[0], p[1] + a.o[1], p[2] + a.o[2]};
    uint32_t c = 0;
    u64* out = emit ? keys + off[i] : nullptr;
    if (isfinite(w[0]) && isfinite(w[1]) && isfinite(w[2]) && w[0] >= 0 && w[1] >= 0 && w[2] >= 0) {
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
                            if (in_box(a.dom[lev], x / rr, y / rr, z / rr, idx)) {
                                if (emit) *out++ = (u64(lev) << 56) | idx;
                                else c++;
                            }
                        } else if (lev == 0 && !emit) {
                            a.marker[id] = 1u;
                        }
                    }
        }
    }
    if (!emit) cnt[i] = c;
}

hipError_t launch_incre_keys(const IncreMarkArgs& a, uint32_t* cnt, const uint32_t* off, u64* keys, bool emit,
                             hipStream_t s) {
    if (a.pts.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_incre_keys, dim3(grid_of(u64(a.pts.n), 256, 1u << 31)), dim3(256), 0, s, a, cnt, off, keys,
                       emit ? 1 : 0);
    return hipGetLastError();
}

// Each level's range of a sorted (level << 56 | cell) list: range[2 * level]
// = first, range[2 * level + 1] = one past the last (both zero beforehand)
__global__ __launch_bounds__(256) void k_level_counts(const u64* k, u64 n, uint32_t* range) {
    const u64 i = u64(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t lev = uint32_t(k[i] >> 56);
    if (i == 0 || uint32_t(k[i - 1] >> 56) != lev) range[2 * lev] = uint32_t(i);
    if (i + 1 == n || uint32_t(k[i + 1] >> 56) != lev) range[2 * lev + 1] = uint32_t(i + 1);
}

hipError_t launch_level_counts(const u64* k, u64 n, uint32_t* counts, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_level_counts, dim3(grid_of(n, 256, ~0ull)), dim3(256), 0, s, k, n, counts);
    return hipGetLastError();
}

// New nodes of one level from its slice of the sorted list: node base + i,
// child list cbase + i (the bitmap form's k_incre_emit)
__global__ __launch_bounds__(256) void k_incre_emit_list(const u64* k, uint32_t n, BitDomain d, Node* nodes, int lev,
                                                         int range, int has_clist, u64 base, u64 cbase) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int x, y, z;
    box_coords(d, k[i] & ((u64(1) << 56) - 1), x, y, z);
    list_node(nodes + base + i, lev, x, y, z, range, has_clist ? Elem(1, lev, cbase + i) : ID_UNDEFL);
}

hipError_t launch_incre_emit_list(const u64* k, uint32_t n, const BitDomain& d, Node* nodes, int lev, int range,
                                  int has_clist, u64 base, u64 cbase, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_incre_emit_list, dim3(grid_of(n, 256, ~0ull)), dim3(256), 0, s, k, n, d, nodes, lev, range,
                       has_clist, base, cbase);
    return hipGetLastError();
}

}  // namespace gvdb
