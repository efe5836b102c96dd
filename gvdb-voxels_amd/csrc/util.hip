// util.hip -- fills and a wave64 exclusive scan (replaces the reference's
// prefixSum/prefixFixup kernels, kernels/cuda_gvdb_particles.cuh:856-902).
#include "kernels.h"

namespace gvdb {

static const int SCAN_THREADS = 256;
static const int SCAN_PER_THREAD = 4;
static const int SCAN_TILE = SCAN_THREADS * SCAN_PER_THREAD;   // 1024 elements per block

__global__ void k_fill_u32(uint32_t* p, uint32_t v, u64 n) {
    u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    u64 stride = u64(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
__global__ void k_fill_u64(u64* p, u64 v, u64 n) {
    u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    u64 stride = u64(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

static unsigned fill_grid(u64 n) {
    u64 b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b == 0) b = 1;
    return unsigned(b);
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, u64 n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}
hipError_t launch_fill_u64(u64* p, u64 v, u64 n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

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
// sums[blockIdx.x] (and to *total when the grid is a single block).
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile(const uint32_t* in0, const uint32_t* in1, int mode,
                                                            uint32_t* out, u64 n, uint32_t* sums, uint32_t* total) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    u64 base = u64(blockIdx.x) * SCAN_TILE + u64(threadIdx.x) * SCAN_PER_THREAD;
    uint32_t v[SCAN_PER_THREAD];
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        u64 i = base + k;
        v[k] = (i < n) ? scan_load(in0, in1, mode, i) : 0u;
        run += v[k];
    }
    // wave64 inclusive scan of the per-thread sums
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
        if (i < n) out[i] = excl;
        excl += v[k];
    }
    if (threadIdx.x == 0) {
        sums[blockIdx.x] = tile;
        if (total && gridDim.x == 1) *total = tile;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_add(uint32_t* out, u64 n, const uint32_t* offs) {
    uint32_t add = offs[blockIdx.x];
    u64 base = u64(blockIdx.x) * SCAN_TILE;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        u64 i = base + u64(k) * SCAN_THREADS + threadIdx.x;
        if (i < n) out[i] += add;
    }
}

u64 scan_tmp_words(u64 n) {
    u64 w = 0;
    do {
        n = (n + SCAN_TILE - 1) / SCAN_TILE;
        w += n + 1;
    } while (n > 1);
    return w + 16;
}

hipError_t launch_scan_u32(const uint32_t* in0, const uint32_t* in1, int mode, uint32_t* out, u64 n,
                           uint32_t* total, uint32_t* tmp, hipStream_t s) {
    if (n == 0) return total ? launch_fill_u32(total, 0u, 1, s) : hipSuccess;
    u64 nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan_tile, dim3(unsigned(nb)), dim3(SCAN_THREADS), 0, s, in0, in1, mode, out, n, tmp, total);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nb == 1) return e;
    // scan the tile sums in place, then add them back
    e = launch_scan_u32(tmp, nullptr, 0, tmp, nb, total, tmp + nb + 1, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3(unsigned(nb)), dim3(SCAN_THREADS), 0, s, out, n, tmp);
    return hipGetLastError();
}

}  // namespace gvdb
