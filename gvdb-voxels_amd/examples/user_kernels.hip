// Sample user kernels for gvdb_compute_kernel / gvdb_render_kernel, built
// into a code object next to the library (lib/user_kernels.hsaco).
#include <hip/hip_runtime.h>
#include "gvdb_amd_device.h"

// v = 2 v + 1 on every brick interior (launched with skipOverAprons).
extern "C" __global__ void user_scale(const gvdb_device_view* v, int3 res, unsigned char chan) {
    int x, y, z;
    if (!gvdb_packed_voxel(v, res, &x, &y, &z)) return;
    const uint64_t i = gvdb_atlas_index(v, x, y, z);
    gvdb_store(v, i, 2.0f * gvdb_load(v, i) + 1.0f);
}

// Each pixel of a float4 render buffer gets (x, y, channel, used bricks).
extern "C" __global__ void user_pixels(const gvdb_device_view* v, unsigned char chan, void* out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= v->width || y >= v->height) return;
    reinterpret_cast<float4*>(out)[y * v->width + x] = make_float4(float(x), float(y), float(chan), float(v->used));
}
