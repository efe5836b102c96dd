// util.hip -- fills, a wave64 exclusive scan (replaces the reference's
// prefixSum/prefixFixup kernels, kernels/cuda_gvdb_particles.cuh:856-902),
// and the staging of argument tables for launches over many grids.
#include "kernels.h"
#include <algorithm>
#include <cstring>

namespace gvdb {

// ------------------------------------------------------- multi-grid setup
hipError_t Arena::init(size_t bytes) {
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&host), bytes, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    e = hipMalloc(reinterpret_cast<void**>(&dev), bytes);
    if (e != hipSuccess) return e;
    cap = bytes;
    used = 0;
    return hipSuccess;
}

void Arena::release() {
    if (host) (void)hipHostFree(host);
    if (dev) (void)hipFree(dev);
    host = dev = nullptr;
    cap = used = 0;
}

void* Arena::push(const void* src, size_t n, hipStream_t s, hipError_t& e) {
    size_t at = (used + 255) & ~size_t(255);
    if (at + n > cap) {
        if (fixed) { e = hipErrorOutOfMemory; return nullptr; }
        // the stream drains before the staging is reused
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return nullptr;
        at = 0;
        if (n > cap) { e = hipErrorOutOfMemory; return nullptr; }
    }
    memcpy(host + at, src, n);
    e = hipMemcpyAsync(dev + at, host + at, n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return nullptr;
    used = at + n;
    return dev + at;
}

hipError_t multi_setup_raw(Arena* ar, const void* args, size_t rec, const uint32_t* nb, int G, hipStream_t s,
                           MultiLaunch& out, bool inline_ok) {
    out.map.G = 1;
    out.map.first = nullptr;
    out.map.table = nullptr;
    out.blocks = 0;
    if (G <= 0) return hipSuccess;
    if (G == 1) {
        out.blocks = nb[0];
        return hipSuccess;
    }
    std::vector<uint32_t> first(size_t(G) + 1);
    u64 acc = 0;
    for (int g = 0; g < G; g++) {
        first[g] = uint32_t(acc);
        acc += nb[g];
    }
    if (acc > 0x7FFFFFFFull) return hipErrorInvalidValue;
    first[G] = uint32_t(acc);
    out.blocks = uint32_t(acc);
    if (acc == 0) return hipSuccess;
    if (inline_ok && G <= MULTI_INLINE && rec * size_t(G) <= size_t(MULTI_INLINE_BYTES)) {
        out.map.G = G;
        memcpy(out.map.ifirst, first.data(), first.size() * 4);
        memcpy(out.map.itab, args, rec * size_t(G));
        return hipSuccess;
    }
    if (!ar) return hipErrorInvalidValue;
    hipError_t e = hipSuccess;
    void* t = ar->push(args, rec * size_t(G), s, e);
    if (!t) return e;
    void* f = ar->push(first.data(), first.size() * 4, s, e);
    if (!f) return e;
    out.map = MultiMap{G, reinterpret_cast<const uint32_t*>(f), t};
    return hipSuccess;
}

// ------------------------------------------------------------------ fills
static const int SCAN_THREADS = 256;
static const int SCAN_PER_THREAD = 4;
static const int SCAN_TILE = SCAN_THREADS * SCAN_PER_THREAD;   // 1024 elements per block

__device__ inline void k_fill_body(const FillArgs& a, uint32_t bid, uint32_t nblk) {
    u64 i = u64(bid) * 256 + threadIdx.x;
    const u64 stride = u64(nblk) * 256;
    if (a.wide) {
        u64* p = reinterpret_cast<u64*>(a.p);
        for (; i < a.n; i += stride) p[i] = a.v;
    } else {
        uint32_t* p = reinterpret_cast<uint32_t*>(a.p);
        const uint32_t v = uint32_t(a.v);
        for (; i < a.n; i += stride) p[i] = v;
    }
}
GVDB_MULTI(k_fill, FillArgs, 256)

static uint32_t fill_grid(u64 n) {
    u64 b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b == 0) b = 1;
    return uint32_t(b);
}

hipError_t launch_fill_multi(const FillArgs* a, int G, Arena* ar, hipStream_t s) {
    std::vector<uint32_t> nb(G);
    for (int g = 0; g < G; g++) nb[g] = a[g].n ? fill_grid(a[g].n) : 0;
    MultiLaunch L;
    hipError_t e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess || L.blocks == 0) return e;
    hipLaunchKernelGGL(k_fill, dim3(L.blocks), dim3(256), 0, s, a[0], L.map);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, u64 n, hipStream_t s) {
    FillArgs a{p, n, v, 0};
    return n ? launch_fill_multi(&a, 1, nullptr, s) : hipSuccess;
}
hipError_t launch_fill_u64(u64* p, u64 v, u64 n, hipStream_t s) {
    FillArgs a{p, n, v, 1};
    return n ? launch_fill_multi(&a, 1, nullptr, s) : hipSuccess;
}

// ------------------------------------------------------------ 64-bit sum
__global__ __launch_bounds__(256) void k_sum_u32(const uint32_t* in, u64 n, u64* out) {
    __shared__ u64 part[4];
    u64 acc = 0;
    for (u64 i = u64(blockIdx.x) * 256 + threadIdx.x; i < n; i += u64(gridDim.x) * 256) acc += in[i];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_down(acc, d, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(reinterpret_cast<unsigned long long*>(out), part[0] + part[1] + part[2] + part[3]);
}

hipError_t launch_sum_u32(const uint32_t* in, u64 n, u64* out, hipStream_t s) {
    hipError_t e = launch_fill_u64(out, 0, 1, s);
    if (e != hipSuccess || n == 0) return e;
    const u64 blocks = std::min<u64>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_sum_u32, dim3(unsigned(blocks)), dim3(256), 0, s, in, n, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------- scan
__device__ inline uint32_t scan_load(const uint32_t* in0, const uint32_t* in1, int mode, u64 i) {
    uint32_t a = in0[i];
    switch (mode) {
        case 1: return __popc(a);
        case 2: return __popc(in1[i] & ~a);
        case 3: return a != 0u ? 1u : 0u;
        default: return a;
    }
}

// Block-local exclusive scan of a 1024-element tile; writes the tile sum to
// sums[tile] (and to *total when the segment is a single tile).
__device__ inline void k_scan_tile_body(const ScanArgs& a, uint32_t bid, uint32_t nblk) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    u64 base = u64(bid) * SCAN_TILE + u64(threadIdx.x) * SCAN_PER_THREAD;
    uint32_t v[SCAN_PER_THREAD];
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        u64 i = base + k;
        v[k] = (i < a.n) ? scan_load(a.in0, a.in1, a.mode, i) : 0u;
        run += v[k];
    }
    int lane = threadIdx.x & 63;
    int wave = threadIdx.x >> 6;
    uint32_t x = run;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t wpre = 0, tile = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; w++) {
        uint32_t s = wsum[w];
        if (w < wave) wpre += s;
        tile += s;
    }
    uint32_t excl = wpre + x - run;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        u64 i = base + k;
        if (i < a.n) a.out[i] = excl;
        excl += v[k];
    }
    if (threadIdx.x == 0) {
        a.sums[bid] = tile;
        if (a.total && nblk == 1) *a.total = tile;
    }
}
GVDB_MULTI(k_scan_tile, ScanArgs, SCAN_THREADS)

struct AddArgs {
    uint32_t* out;
    u64 n;
    const uint32_t* offs;
};
__device__ inline void k_scan_add_body(const AddArgs& a, uint32_t bid, uint32_t) {
    uint32_t add = a.offs[bid];
    u64 base = u64(bid) * SCAN_TILE;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        u64 i = base + u64(k) * SCAN_THREADS + threadIdx.x;
        if (i < a.n) a.out[i] += add;
    }
}
GVDB_MULTI(k_scan_add, AddArgs, SCAN_THREADS)

u64 scan_tmp_words(u64 n) {
    u64 w = 0;
    do {
        n = (n + SCAN_TILE - 1) / SCAN_TILE;
        w += n + 1;
    } while (n > 1);
    return w + 16;
}

hipError_t launch_scan_multi(const ScanArgs* a, int G, Arena* ar, hipStream_t s) {
    if (G <= 0) return hipSuccess;
    std::vector<uint32_t> nb(G);
    std::vector<FillArgs> zero;
    for (int g = 0; g < G; g++) {
        nb[g] = uint32_t((a[g].n + SCAN_TILE - 1) / SCAN_TILE);
        if (a[g].n == 0 && a[g].total) zero.push_back(FillArgs{a[g].total, 1, 0, 0});
    }
    hipError_t e = hipSuccess;
    if (!zero.empty()) {
        e = launch_fill_multi(zero.data(), int(zero.size()), ar, s);
        if (e != hipSuccess) return e;
    }
    MultiLaunch L;
    e = multi_setup(ar, a, nb.data(), G, s, L);
    if (e != hipSuccess) return e;
    if (L.blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scan_tile, dim3(L.blocks), dim3(SCAN_THREADS), 0, s, a[0], L.map);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // segments of more than one tile: scan their tile sums in place, add back
    std::vector<ScanArgs> up;
    std::vector<AddArgs> add;
    std::vector<uint32_t> addnb;
    for (int g = 0; g < G; g++) {
        if (nb[g] <= 1) continue;
        up.push_back(ScanArgs{a[g].sums, nullptr, 0, a[g].sums, nb[g], a[g].sums + nb[g] + 1, a[g].total});
        add.push_back(AddArgs{a[g].out, a[g].n, a[g].sums});
        addnb.push_back(nb[g]);
    }
    if (up.empty()) return hipSuccess;
    e = launch_scan_multi(up.data(), int(up.size()), ar, s);
    if (e != hipSuccess) return e;
    MultiLaunch L2;
    e = multi_setup(ar, add.data(), addnb.data(), int(add.size()), s, L2);
    if (e != hipSuccess || L2.blocks == 0) return e;
    hipLaunchKernelGGL(k_scan_add, dim3(L2.blocks), dim3(SCAN_THREADS), 0, s, add[0], L2.map);
    return hipGetLastError();
}

hipError_t launch_scan_u32(const uint32_t* in0, const uint32_t* in1, int mode, uint32_t* out, u64 n,
                           uint32_t* total, uint32_t* tmp, hipStream_t s) {
    ScanArgs a{in0, in1, mode, out, n, tmp, total};
    return launch_scan_multi(&a, 1, nullptr, s);
}

}  // namespace gvdb
