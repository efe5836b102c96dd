// compute.hip -- Compute(effect, ...) operators on float and half atlas
// channels (src/gvdb_volume_gvdb.cpp:4527-4552; kernels/cuda_gvdb_operators.cuh
// :148-455 of the reference).
//
// The reference launches 8x8x8 blocks over the 3D atlas texture.  Its
// "packed" operators (smooth, noise, grow) give each block one atlas brick:
// the block reads the brick's padded 10^3 voxels into shared memory and
// writes the 8^3 interior, so bricks never read each other's writes.  Here
// one 512-thread workgroup does the same for one brick of the brick-major
// atlas, over the bricks that hold leaves.  The "unpacked" fill writes every
// atlas voxel of [1, min(8 * blocks, res) - 1] per axis, apron voxels
// included, as the reference's indexing does.
#include "kernels.h"
#include <hip/hip_fp16.h>
#include <algorithm>

#pragma clang fp contract(off)

namespace gvdb {

namespace {
template <typename AT> __device__ inline float ldv(const void* a, u64 o);
template <> __device__ inline float ldv<float>(const void* a, u64 o) { return static_cast<const float*>(a)[o]; }
template <> __device__ inline float ldv<__half>(const void* a, u64 o) {
    return __half2float(static_cast<const __half*>(a)[o]);
}
template <typename AT> __device__ inline void stv(void* a, u64 o, float v);
template <> __device__ inline void stv<float>(void* a, u64 o, float v) { static_cast<float*>(a)[o] = v; }
template <> __device__ inline void stv<__half>(void* a, u64 o, float v) { static_cast<__half*>(a)[o] = __float2half(v); }

// Bob Jenkins' one-at-a-time final mix and its compound forms, and the
// [0, 1) float built from the low 23 bits (random(float3),
// kernels/cuda_gvdb.cuh:135-164).
__device__ inline uint32_t oaat(uint32_t x) {
    x += (x << 10u);
    x ^= (x >> 6u);
    x += (x << 3u);
    x ^= (x >> 11u);
    x += (x << 15u);
    return x;
}
__device__ inline float random3(float a, float b, float c) {
    const uint32_t h = oaat(uint32_t(__float_as_int(a)) ^ oaat(uint32_t(__float_as_int(b))) ^
                            oaat(uint32_t(__float_as_int(c))));
    return __int_as_float(int((h & 0x007FFFFFu) | 0x3F800000u)) - 1.0f;
}
}  // namespace

// One workgroup per brick slot: thread t owns interior voxel (t & 7, t >> 3 & 7,
// t >> 6) + 1 of the padded brick.
template <typename AT>
__global__ __launch_bounds__(512) void k_compute_packed(ComputeArgs a) {
    const u64 slot = blockIdx.x;
    const int br = a.t.brick_res;
    const int t = threadIdx.x;
    const int lx = (t & 7) + 1, ly = ((t >> 3) & 7) + 1, lz = (t >> 6) + 1;
    const u64 base = slot * u64(a.t.brick_vox);
    auto at = [&](int x, int y, int z) { return base + u64((z * br + y) * br + x); };
    const float c = ldv<AT>(a.atlas, at(lx, ly, lz));
    float v = c;
    if (a.op == OP_SMOOTH) {
        // v = (p1 c + the six face neighbours) / (p1 + 6.0) + p2, the division
        // and the sum with p2 in double as the reference's double literal makes them
        v = a.p1 * c;
        v += ldv<AT>(a.atlas, at(lx - 1, ly, lz));
        v += ldv<AT>(a.atlas, at(lx + 1, ly, lz));
        v += ldv<AT>(a.atlas, at(lx, ly - 1, lz));
        v += ldv<AT>(a.atlas, at(lx, ly + 1, lz));
        v += ldv<AT>(a.atlas, at(lx, ly, lz - 1));
        v += ldv<AT>(a.atlas, at(lx, ly, lz + 1));
        v = float(double(v) / (double(a.p1) + 6.0) + double(a.p2));
    } else if (a.op == OP_NOISE) {
        if (double(v) > 0.01) {
            const int sx = int(slot % u64(a.t.atlas_cnt[0]));
            const int sy = int((slot / u64(a.t.atlas_cnt[0])) % u64(a.t.atlas_cnt[1]));
            const int sz = int(slot / (u64(a.t.atlas_cnt[0]) * a.t.atlas_cnt[1]));
            v += random3(float(sx * br + lx), float(sy * br + ly), float(sz * br + lz)) * a.p1;
        }
    } else if (a.op == OP_GROW) {
        if (v != 0.0f) v = float(double(v) + double(a.p1) * 10.0);
        if (double(v) < 0.01) v = 0.0f;
    }
    __syncthreads();   // every read of the brick before any write
    stv<AT>(a.atlas, at(lx, ly, lz), v);
}

// Atlas voxels [lo, hi] (inclusive) per axis, in atlas coordinates.
template <typename AT>
__global__ __launch_bounds__(256) void k_compute_fill(ComputeArgs a) {
    const u64 nx = u64(a.hi[0] - a.lo[0] + 1), ny = u64(a.hi[1] - a.lo[1] + 1), nz = u64(a.hi[2] - a.lo[2] + 1);
    const u64 n = nx * ny * nz;
    for (u64 i = u64(blockIdx.x) * 256 + threadIdx.x; i < n; i += u64(gridDim.x) * 256) {
        const int x = a.lo[0] + int(i % nx), y = a.lo[1] + int((i / nx) % ny), z = a.lo[2] + int(i / (nx * ny));
        float v = a.p1;
        if (a.p3 < 0) {
            v = sinf(float(double(uint32_t(x) * 12u) / (3.141592 * 30.0)));
            v += sinf(float(double(uint32_t(y) * 12u) / (3.141592 * 30.0)));
            v += sinf(float(double(uint32_t(z) * 12u) / (3.141592 * 30.0)));
        }
        stv<AT>(a.atlas, atlas_offset(a.t, x, y, z), v);
    }
}

hipError_t launch_compute(const ComputeArgs& a, hipStream_t s) {
    const bool half = a.atlas_type == T_HALF;
    if (a.op == OP_FILL) {
        for (int k = 0; k < 3; k++)
            if (a.hi[k] < a.lo[k]) return hipSuccess;
        const u64 n = u64(a.hi[0] - a.lo[0] + 1) * u64(a.hi[1] - a.lo[1] + 1) * u64(a.hi[2] - a.lo[2] + 1);
        const unsigned g = unsigned(std::min<u64>((n + 255) / 256, 65536));
        if (half) hipLaunchKernelGGL(k_compute_fill<__half>, dim3(g), dim3(256), 0, s, a);
        else hipLaunchKernelGGL(k_compute_fill<float>, dim3(g), dim3(256), 0, s, a);
        return hipGetLastError();
    }
    if (a.slots <= 0) return hipSuccess;
    if (a.t.brickwid != 8 || a.t.brick_res != 10) return hipErrorInvalidValue;
    if (half) hipLaunchKernelGGL(k_compute_packed<__half>, dim3(unsigned(a.slots)), dim3(512), 0, s, a);
    else hipLaunchKernelGGL(k_compute_packed<float>, dim3(unsigned(a.slots)), dim3(512), 0, s, a);
    return hipGetLastError();
}

}  // namespace gvdb
