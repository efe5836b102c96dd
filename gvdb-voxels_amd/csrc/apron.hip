// apron.hip -- UpdateApron as neighbour-face copies.
//
// The reference (gvdbUpdateApronF, kernels/cuda_gvdb_operators.cuh:72-131)
// maps every apron voxel of a brick back to index space and walks the tree
// from the top node to find the brick that holds it, reading that brick's
// interior voxel or writing the boundary value when no brick covers it.  All
// apron voxels in one of the 26 directions around a brick fall inside the same
// neighbour brick, so here each brick does the 26 walks once (k_neighbors),
// and the copy kernel moves the 488 apron voxels of a brick from the
// neighbours' interior faces, edges and corners.
#include "kernels.h"
#include "tree_device.h"
#include <hip/hip_fp16.h>
#include <mutex>
#include <cstring>
#include <algorithm>

namespace gvdb {

__device__ inline void dir_of(int d, int& dx, int& dy, int& dz) {
    // d in [0,27) without the centre 13: dx fastest
    dx = d % 3 - 1;
    dy = (d / 3) % 3 - 1;
    dz = d / 9 - 1;
}

// 32 lanes per brick (27 directions used), two bricks per wave: the mask of
// present neighbours comes from one ballot, written by the brick's lane 0.
__device__ inline void k_neighbors_body(const NbrArgs& a, uint32_t bid, uint32_t) {
    const Tree& t = a.t;
    int* nbr = a.nbr;
    const u64 g = u64(bid) * 256 + threadIdx.x;
    const int b = int(g >> 5), d = int(g & 31);
    const bool valid = b < a.n && d < 27;
    int v = -1;
    if (valid) {
        const Node* n = t.nodes[0] + b;
        int dx, dy, dz;
        dir_of(d, dx, dy, dz);
        if (a.bt.tab) {
            // the brick table at the brick's coordinates plus the direction
            // (d = 13: the brick itself, -1 when the table does not hold it)
            const int sh = __builtin_ctz(unsigned(t.brickwid));
            v = tab_leaf(a.bt, (n->mPos[0] >> sh) + dx, (n->mPos[1] >> sh) + dy, (n->mPos[2] >> sh) + dz);
            if (d == 13 && v != b) v = -1;
        } else if (d == 13) {
            // the brick itself, or -1 when the walk from the top node does not reach it
            v = pos_leaf(t, n->mPos[0], n->mPos[1], n->mPos[2]) == b ? b : -1;
        } else {
            int w = t.brickwid;
            // a voxel centre inside the neighbour box: local coordinate -1, 0 or w
            float px = float(n->mPos[0]) + float(dx < 0 ? -1 : (dx > 0 ? w : 0)) + 0.5f;
            float py = float(n->mPos[1]) + float(dy < 0 ? -1 : (dy > 0 ? w : 0)) + 0.5f;
            float pz = float(n->mPos[2]) + float(dz < 0 ? -1 : (dz > 0 ? w : 0)) + 0.5f;
            v = point_leaf(t, px, py, pz);
        }
        const u64 e = u64(b) * 27 + u64(d);
        nbr[e] = v;
        if (a.base) {
            // atlas offset of the neighbour's padded brick (d = 13: the brick's own)
            const int m = d == 13 ? b : v;
            const Node* mn = t.nodes[0] + (m >= 0 ? m : 0);
            const bool live = m >= 0 && (d != 13 || mn->mFlags) && mn->mValue[0] >= 0;
            a.base[e] = live ? (long long)(slot_of(t, mn->mValue) * u64(t.brick_vox)) : -1ll;
        }
    }
    const u64 bal = __ballot(valid && v >= 0);
    if (a.mask && valid && d == 0) {
        const uint32_t m = uint32_t(bal >> (threadIdx.x & 32)) & 0x7FFFFFFu;
        a.mask[b] = m;
        if (a.bt.tab) {
            // beside the brick's index in the table (its own entry: v is b here
            // when the table holds the brick), for the insert's one-read lookup
            const Node* n = t.nodes[0] + b;
            const int sh = __builtin_ctz(unsigned(t.brickwid));
            const unsigned x = unsigned((n->mPos[0] >> sh) - a.bt.lo[0]), y = unsigned((n->mPos[1] >> sh) - a.bt.lo[1]),
                           z = unsigned((n->mPos[2] >> sh) - a.bt.lo[2]);
            if (x < unsigned(a.bt.dim[0]) && y < unsigned(a.bt.dim[1]) && z < unsigned(a.bt.dim[2]))
                reinterpret_cast<uint32_t*>(a.bt.tab)[2 * tab_index(a.bt, x, y, z) + 1] = m;
        }
    }
}

GVDB_MULTI(k_neighbors, NbrArgs, 256)

hipError_t launch_neighbors_multi(const NbrArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> nb(G);
    for (int g = 0; g < G; g++) nb[g] = a[g].n > 0 ? uint32_t((u64(a[g].n) * 32 + 255) / 256) : 0;
    MultiLaunch L;
    hipError_t e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    L.map.xcd = knob(KNOB_XCD_ORDER);
    hipLaunchKernelGGL(k_neighbors, dim3(L.blocks), dim3(256), 0, s, a[0], L.map);
    return hipGetLastError();
}

template <typename AT>
__device__ inline float ld(const AT* a, u64 o);
template <>
__device__ inline float ld<float>(const float* a, u64 o) { return a[o]; }
template <>
__device__ inline float ld<__half>(const __half* a, u64 o) { return __half2float(a[o]); }
template <typename AT>
__device__ inline void st(AT* a, u64 o, float v);
template <>
__device__ inline void st<float>(float* a, u64 o, float v) { a[o] = v; }
template <>
__device__ inline void st<__half>(__half* a, u64 o, float v) { a[o] = __float2half(v); }
// uchar4 channels move as 32-bit words; the value travels through the body as
// a float holding those bits (moved, never computed with), and the boundary
// value arrives already packed (apron_bound)
template <>
__device__ inline float ld<uint32_t>(const uint32_t* a, u64 o) { return __uint_as_float(a[o]); }
template <>
__device__ inline void st<uint32_t>(uint32_t* a, u64 o, float v) { a[o] = __float_as_uint(v); }

// UpdateApron<uchar4>'s boundary value, make_uchar4(boundval, ...) (kernels/
// cuda_gvdb_operators.cuh:143-146): each byte is boundval converted to an
// unsigned char, truncated and clamped to [0, 255] (NaN: 0), packed into the
// bits of a float
static float apron_bound_u8x4(float b) {
    const float c = b > 0.f ? (b < 255.f ? b : 255.f) : 0.f;   // NaN fails b > 0
    const uint32_t u = uint32_t(c) * 0x01010101u;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// Apron voxel a of a padded brick of side br = w + 2: the two z faces
// (br^2 each), then per interior z layer its ring (br^2 - w^2: the y = 0 and
// y = br - 1 rows, then x = 0 / x = br - 1 of each row between).  W > 0: the
// brick width known when compiled (its divisions become multiplies).
template <int W>
__device__ inline void apron_decode(int a, int w_rt, int& x, int& y, int& z) {
    const unsigned w = W > 0 ? unsigned(W) : unsigned(w_rt);
    const unsigned br = w + 2, face = br * br, ring = br * br - w * w, u = unsigned(a);
    if (u < 2 * face) {
        z = u < face ? 0 : int(br) - 1;
        const unsigned r = u < face ? u : u - face;
        y = int(r / br);
        x = int(r - unsigned(y) * br);
    } else {
        const unsigned s = u - 2 * face;
        z = int(s / ring) + 1;
        const int r = int(s - unsigned(z - 1) * ring);
        if (r < int(br)) { x = r; y = 0; }
        else if (r < 2 * int(br)) { x = r - int(br); y = int(br) - 1; }
        else { const int q = r - 2 * int(br); y = (q >> 1) + 1; x = (q & 1) ? int(br) - 1 : 0; }
    }
}

// Apron voxel `a` (0..487) of brick b, given the atlas offsets of the brick
// and its 26 neighbours (-1: none).  Enumeration: the z = 0 and z = 9 faces
// (100 each), then the 36-voxel rings of z = 1..8.
template <typename AT>
__device__ inline void apron_voxel(const ApronArgs& ap, int b, int a, const long long* base) {
    const Tree& t = ap.t;
    AT* atlas = reinterpret_cast<AT*>(ap.atlas);
    const int br = t.brick_res;        // 10
    const int w = t.brickwid;          // 8
    int x, y, z;
    if (w == 8) apron_decode<8>(a, w, x, y, z);   // uniform: the usual 8^3 bricks
    else apron_decode<0>(a, w, x, y, z);
    int dx = x == 0 ? -1 : (x == br - 1 ? 1 : 0);
    int dy = y == 0 ? -1 : (y == br - 1 ? 1 : 0);
    int dz = z == 0 ? -1 : (z == br - 1 ? 1 : 0);
    const int d = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
    const long long nbase = base[d];
    const u64 own = u64(base[13]) + u64((z * br + y) * br + x);
    float v = ap.boundval;
    if (nbase >= 0) {
        // interior coordinate inside the neighbour
        const int sx = (x - 1) - dx * w, sy = (y - 1) - dy * w, sz = (z - 1) - dz * w;
        if (ap.faces) {
            // it lies on the neighbour's face towards this brick (x face first)
            const int f = dx ? (dx < 0 ? 1 : 0) : (dy ? (dy < 0 ? 3 : 2) : (dz < 0 ? 5 : 4));
            const int i = dx ? sz * 8 + sy : (dy ? sz * 8 + sx : sy * 8 + sx);
            const int nb = ap.nbr[u64(b) * 27 + d];
            v = ld<AT>(reinterpret_cast<const AT*>(ap.faces), u64(nb) * FACE_VALS + u64(f * 64 + i));
        } else {
            v = ld<AT>(atlas, u64(nbase) + u64(((sz + 1) * br + sy + 1) * br + sx + 1));
        }
    }
    st<AT>(atlas, own, v);
}

// One 512-thread workgroup per brick; threads 0..487 each own one apron voxel
// of the padded 10^3 brick (apron 1, 8^3 interior).
template <typename AT>
__device__ inline void k_apron_body(const ApronArgs& ap, uint32_t bid, uint32_t) {
    __shared__ long long base[27];
    const Tree& t = ap.t;
    const int b = int(bid);
    const int a = threadIdx.x;
    const int br = t.brick_res, w = t.brickwid;
    // atlas offsets of the brick and its 26 neighbours (-1: none), one load
    if (a < 27) base[a] = ap.base[u64(b) * 27 + a];
    __syncthreads();
    if (a >= 2 * br * br + w * (br * br - w * w) || base[13] < 0) return;
    apron_voxel<AT>(ap, b, a, base);
}

GVDB_MULTI_T(k_apron, typename AT, AT, ApronArgs, 512)

// A run of up to RUN consecutive bricks b0.. of one grid, its voxels spread
// over all 512 threads: the atlas offsets of the run's bricks in one load and
// one barrier, then RUN * 488 voxels in turn (a workgroup per brick launched
// eight waves for 488 one-voxel threads).  Ends with a barrier, so the next
// run may reuse the offsets' LDS.
template <typename AT, int RUN>
__device__ inline void apron_run(const ApronArgs& a0, uint32_t b0, long long* base) {
    const int nb = min(RUN, int(uint32_t(a0.n) - b0));
    for (int i = threadIdx.x; i < nb * 27; i += 512) base[i] = a0.base[u64(b0) * 27 + u64(i)];
    __syncthreads();
    const int br = a0.t.brick_res, w = a0.t.brickwid;
    const int per = 2 * br * br + w * (br * br - w * w);
    for (int v = threadIdx.x; v < nb * per; v += 512) {
        const int k = w == 8 ? int(unsigned(v) / 488u) : v / per, a = v - k * per;   // 488 = per for 8^3 bricks
        if (base[k * 27 + 13] < 0) continue;
        apron_voxel<AT>(a0, int(b0) + k, a, base + k * 27);
    }
    __syncthreads();
}

// Several grids: RUN bricks of one grid per workgroup (the apron-run knob:
// 1, 2, 4 or 8), one after the other, so the grid lookup and the record
// loads before the first brick are paid once per run (a brick alone is one
// voxel per thread).
template <typename AT, int RUN>
__global__ __launch_bounds__(512) void k_apron_runs(ApronArgs, MultiMap m) {
    __shared__ long long base[RUN * 27];
    const uint32_t bi = m.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const int g = multi_grid(m, bi);
    const uint32_t* first = multi_first(m);
    const ApronArgs& a = multi_rec<ApronArgs>(m, g);
    const uint32_t b0 = (bi - first[g]) * uint32_t(RUN);
    if (b0 < uint32_t(a.n)) apron_run<AT, RUN>(a, b0, base);
}

// Several grids, persistent: a fixed number of workgroups (about as many as
// the device holds at once) take runs of QRUN bricks from one device counter
// until the queue is empty.  A run's grid is looked up only when it falls
// outside the workgroup's previous grid; every workgroup leaves when the
// counter passes the last run.
static const int QRUN = 8;
static const uint32_t APRON_QUEUE_WORKGROUPS = 256 * 4;   // 4 per CU of 256
template <typename AT>
__global__ __launch_bounds__(512) void k_apron_queue(ApronArgs, MultiMap m, uint32_t* counter, uint32_t nruns) {
    __shared__ uint32_t s_run;
    __shared__ long long base[QRUN * 27];
    const uint32_t* first = multi_first(m);
    int g = -1;
    uint32_t lo = 0, hi = 0;
    for (;;) {
        if (threadIdx.x == 0) s_run = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t run = s_run;
        __syncthreads();                   // read before thread 0 takes the next run
        if (run >= nruns) break;
        if (g < 0 || run < lo || run >= hi) {
            g = multi_grid(m, run);
            lo = first[g];
            hi = first[g + 1];
        }
        const ApronArgs& a = multi_rec<ApronArgs>(m, g);
        const uint32_t b0 = (run - lo) * uint32_t(QRUN);
        if (b0 < uint32_t(a.n)) apron_run<AT, QRUN>(a, b0, base);
    }
}

template <typename AT>
static void launch_apron_runs(int run, dim3 grid, hipStream_t s, const ApronArgs& a0, const MultiMap& m) {
    switch (run) {
    case 1: hipLaunchKernelGGL((k_apron_runs<AT, 1>), grid, dim3(512), 0, s, a0, m); break;
    case 2: hipLaunchKernelGGL((k_apron_runs<AT, 2>), grid, dim3(512), 0, s, a0, m); break;
    case 8: hipLaunchKernelGGL((k_apron_runs<AT, 8>), grid, dim3(512), 0, s, a0, m); break;
    default: hipLaunchKernelGGL((k_apron_runs<AT, 4>), grid, dim3(512), 0, s, a0, m); break;
    }
}

// One grid: runs of RUN consecutive bricks per workgroup, the record by value.
template <typename AT, int RUN>
__global__ __launch_bounds__(512) void k_apron_flat(ApronArgs a0, MultiMap m) {
    __shared__ long long base[RUN * 27];
    const uint32_t bi = m.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t b0 = bi * uint32_t(RUN);
    if (b0 < uint32_t(a0.n)) apron_run<AT, RUN>(a0, b0, base);
}

template <typename AT>
static void launch_apron_seq(int run, dim3 grid, hipStream_t s, const ApronArgs& a0, const MultiMap& m) {
    switch (run) {
    case 2: hipLaunchKernelGGL((k_apron_flat<AT, 2>), grid, dim3(512), 0, s, a0, m); break;
    case 4: hipLaunchKernelGGL((k_apron_flat<AT, 4>), grid, dim3(512), 0, s, a0, m); break;
    case 8: hipLaunchKernelGGL((k_apron_flat<AT, 8>), grid, dim3(512), 0, s, a0, m); break;
    case 16: hipLaunchKernelGGL((k_apron_flat<AT, 16>), grid, dim3(512), 0, s, a0, m); break;
    case 32: hipLaunchKernelGGL((k_apron_flat<AT, 32>), grid, dim3(512), 0, s, a0, m); break;
    default: hipLaunchKernelGGL(k_apron<AT>, grid, dim3(512), 0, s, a0, m); break;
    }
}

template <typename AT>
static void launch_apron_as(bool queue, bool runs, int run, uint32_t nwg, uint32_t* ctr, const MultiLaunch& L,
                            hipStream_t s, const ApronArgs& a0) {
    if (queue)
        hipLaunchKernelGGL(k_apron_queue<AT>, dim3(nwg), dim3(512), 0, s, a0, L.map, ctr, L.blocks);
    else if (runs) launch_apron_runs<AT>(run, dim3(L.blocks), s, a0, L.map);
    else launch_apron_seq<AT>(run, dim3(L.blocks), s, a0, L.map);
}

hipError_t launch_apron_multi(const ApronArgs* a, int G, Arena* ar, hipStream_t s) {
    // one launch per channel type: 0 float, 1 half, 2 uchar4
    for (int kind = 0; kind < 3; kind++) {
        std::vector<ApronArgs> sel;
        std::vector<uint32_t> nb;
        for (int g = 0; g < G; g++) {
            if (a[g].t.apron != 1) return hipErrorInvalidValue;
            const int k = a[g].atlas_type == T_HALF ? 1 : (a[g].atlas_type == T_UCHAR4 ? 2 : 0);
            if (k != kind || a[g].n <= 0) continue;
            sel.push_back(a[g]);
            if (kind == 2) {
                sel.back().boundval = apron_bound_u8x4(a[g].boundval);
                sel.back().faces = nullptr;
            }
        }
        if (sel.empty()) continue;
        // one grid: runs of `run` consecutive bricks per workgroup (1: a
        // workgroup per brick); several: runs of one grid's bricks, or the
        // persistent queue over runs of QRUN bricks
        const bool runs = sel.size() > 1;
        const bool queue = runs && knob(KNOB_APRON_QUEUE) == 1;
        const int run = queue ? QRUN : (runs ? knob(KNOB_APRON_RUN) : knob(KNOB_APRON_RUN_ONE));
        for (const ApronArgs& x : sel) nb.push_back(uint32_t((x.n + run - 1) / run));
        MultiLaunch L;
        hipError_t e = multi_setup(ar, sel.data(), nb.data(), int(sel.size()), s, L);
        if (e != hipSuccess) return e;
        if (L.blocks == 0) continue;
        L.map.xcd = knob(KNOB_XCD_ORDER);
        uint32_t* ctr = nullptr;
        const uint32_t nwg = std::min<uint32_t>(L.blocks, APRON_QUEUE_WORKGROUPS);
        if (queue) {
            // the counter starts at zero: a copy recorded on the stream (also
            // inside a graph), from the argument staging
            const uint32_t zero = 0;
            ctr = ar ? static_cast<uint32_t*>(ar->push(&zero, sizeof(zero), s, e)) : nullptr;
            if (!ctr) return e != hipSuccess ? e : hipErrorInvalidValue;
        }
        if (kind == 1) launch_apron_as<__half>(queue, runs, run, nwg, ctr, L, s, sel[0]);
        else if (kind == 2) launch_apron_as<uint32_t>(queue, runs, run, nwg, ctr, L, s, sel[0]);
        else launch_apron_as<float>(queue, runs, run, nwg, ctr, L, s, sel[0]);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ------------------------------------------- UpdateApron without allocation
// The neighbour table of gvdb_update_apron_noalloc.  Every device that loads
// this code object holds its own copy until the process ends.  In host code
// the name stands for no device address: noalloc_scratch() looks up the
// current device's copy with hipGetSymbolAddress, and only that address is
// handed to kernels.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static const int NOALLOC_DEVICES = 64;

hipError_t noalloc_scratch(int device, void** ptr) {
    static std::mutex mu;
    static void* addr[NOALLOC_DEVICES] = {};
    *ptr = nullptr;
    if (device < 0 || device >= NOALLOC_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *ptr = addr[device];
    return hipSuccess;
}

hipError_t launch_apron_noalloc(const Tree& t, void* sp, const void* faces, void* atlas, int atlas_type,
                                float boundval, int bricks, hipStream_t s) {
    if (bricks <= 0) return hipSuccess;
    if (!sp || bricks > NOALLOC_MAX_BRICKS || t.apron != 1) return hipErrorInvalidValue;
    // offsets first (8-byte entries), then the leaf indices
    long long* base = static_cast<long long*>(sp);
    int* nbr = reinterpret_cast<int*>(static_cast<char*>(sp) + size_t(bricks) * 27 * 8);
    const MultiMap one{1, nullptr, nullptr, knob(KNOB_XCD_ORDER), {}, {}};
    // every entry of both tables is written by the neighbour pass before the
    // copy reads it, so nothing from an earlier call is read
    const NbrArgs na{t, nbr, bricks, nullptr, base};
    hipLaunchKernelGGL(k_neighbors, dim3(uint32_t((u64(bricks) * 32 + 255) / 256)), dim3(256), 0, s, na, one);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (atlas_type == T_UCHAR4) {
        const ApronArgs aa{t, base, nbr, nullptr, atlas, atlas_type, apron_bound_u8x4(boundval), bricks};
        hipLaunchKernelGGL(k_apron<uint32_t>, dim3(uint32_t(bricks)), dim3(512), 0, s, aa, one);
        return hipGetLastError();
    }
    const ApronArgs aa{t, base, nbr, faces, atlas, atlas_type, boundval, bricks};
    if (atlas_type == T_HALF) hipLaunchKernelGGL(k_apron<__half>, dim3(uint32_t(bricks)), dim3(512), 0, s, aa, one);
    else hipLaunchKernelGGL(k_apron<float>, dim3(uint32_t(bricks)), dim3(512), 0, s, aa, one);
    return hipGetLastError();
}

}  // namespace gvdb
