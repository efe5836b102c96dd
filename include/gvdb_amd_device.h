/* gvdb_amd_device.h -- what a user kernel launched by gvdb_compute_kernel or
 * gvdb_render_kernel receives, and inline helpers to address the atlas.
 *
 * Replaces the reference's VDBInfo/surface access for user kernels
 * (ComputeKernel, RenderKernel: gvdb_volume_gvdb.cpp:4309-4333, 4499-4525;
 * kernels/cuda_gvdb_operators.cuh GVDB_VOXPACKED/GVDB_VOXUNPACKED).
 *
 * Compute kernels are launched as the reference launches them: 8x8x8 blocks,
 * ceil(count / 8) blocks per axis, count = the atlas resolution, or the
 * interior-only ("packed") resolution when skipOverAprons is set, with
 *     extern "C" __global__ void k(const gvdb_device_view* v, int3 atlas_res, unsigned char chan);
 * Render kernels get 8x8x1 blocks over the render buffer with
 *     extern "C" __global__ void k(const gvdb_device_view* v, unsigned char chan, void* out);
 *
 * The atlas is brick-major: voxel (x, y, z) of the atlas (the reference's 3D
 * texture coordinates) lives at element gvdb_atlas_index(v, x, y, z) of
 * v->atlas.  Include from HIP device code. */
#ifndef GVDB_AMD_DEVICE_H
#define GVDB_AMD_DEVICE_H

#include <stdint.h>

#define GVDB_VIEW_MAXLEV 10

typedef struct gvdb_device_view {
    void* atlas;                 /* the channel's data */
    int type;                    /* GVDB_T_FLOAT (3) or GVDB_T_HALF (100) */
    int brick_res;               /* padded brick side: brickwid + 2 * apron */
    int brickwid;
    int apron;
    int atlas_cnt[3];            /* bricks per axis */
    int atlas_res[3];            /* atlas_cnt * brick_res */
    int used;                    /* atlas slots holding a leaf: 0 .. used - 1 */
    const int32_t* amap;         /* per slot: corner x, y, z of its leaf, leaf index (-1: none) */
    /* tree: node pools (64-byte nodes) and child lists per level */
    int levels, top_lev;
    int res[GVDB_VIEW_MAXLEV], range[GVDB_VIEW_MAXLEV], nodecnt[GVDB_VIEW_MAXLEV];
    const void* nodes[GVDB_VIEW_MAXLEV];
    const uint64_t* clists[GVDB_VIEW_MAXLEV];
    /* render kernels: the scene block (68 words, gvdb_get_scene_block's layout) */
    const float* scene;
    int width, height;
} gvdb_device_view;

#if defined(__HIPCC__)
#include <hip/hip_fp16.h>

static __device__ inline uint64_t gvdb_atlas_index(const gvdb_device_view* v, int x, int y, int z) {
    const int br = v->brick_res;
    const int sx = x / br, sy = y / br, sz = z / br;
    const uint64_t slot = ((uint64_t)sz * v->atlas_cnt[1] + sy) * v->atlas_cnt[0] + sx;
    return slot * (uint64_t)(br * br * br) + (uint64_t)(((z - sz * br) * br + (y - sy * br)) * br + (x - sx * br));
}

/* GVDB_VOXPACKED: the atlas voxel of this thread when blocks walk brick
 * interiors (skipOverAprons); returns 0 outside the atlas. */
static __device__ inline int gvdb_packed_voxel(const gvdb_device_view* v, int3 res, int* x, int* y, int* z) {
    const int w = v->brick_res - 2 * v->apron;
    const int px = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y * blockDim.y + threadIdx.y,
              pz = blockIdx.z * blockDim.z + threadIdx.z;
    *x = px + (px / w) * 2 * v->apron + v->apron;
    *y = py + (py / w) * 2 * v->apron + v->apron;
    *z = pz + (pz / w) * 2 * v->apron + v->apron;
    return *x < res.x && *y < res.y && *z < res.z;
}

static __device__ inline float gvdb_load(const gvdb_device_view* v, uint64_t i) {
    if (v->type == 100) return __half2float(((const __half*)v->atlas)[i]);
    return ((const float*)v->atlas)[i];
}
static __device__ inline void gvdb_store(const gvdb_device_view* v, uint64_t i, float val) {
    if (v->type == 100) ((__half*)v->atlas)[i] = __float2half(val);
    else ((float*)v->atlas)[i] = val;
}
#endif

#endif
