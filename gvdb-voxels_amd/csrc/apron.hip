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

namespace gvdb {

__device__ inline void dir_of(int d, int& dx, int& dy, int& dz) {
    // d in [0,27) without the centre 13: dx fastest
    dx = d % 3 - 1;
    dy = (d / 3) % 3 - 1;
    dz = d / 9 - 1;
}

__global__ __launch_bounds__(256) void k_neighbors(Tree t, int* nbr, int nbricks) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nbricks * 27) return;
    int b = g / 27, d = g - b * 27;
    const Node* n = t.nodes[0] + b;
    if (d == 13) {
        // the brick itself, or -1 when the walk from the top node does not reach it
        nbr[g] = pos_leaf(t, n->mPos[0], n->mPos[1], n->mPos[2]) == b ? b : -1;
        return;
    }
    int dx, dy, dz;
    dir_of(d, dx, dy, dz);
    int w = t.brickwid;
    // a voxel centre inside the neighbour box: local coordinate -1, 0 or w
    float px = float(n->mPos[0]) + float(dx < 0 ? -1 : (dx > 0 ? w : 0)) + 0.5f;
    float py = float(n->mPos[1]) + float(dy < 0 ? -1 : (dy > 0 ? w : 0)) + 0.5f;
    float pz = float(n->mPos[2]) + float(dz < 0 ? -1 : (dz > 0 ? w : 0)) + 0.5f;
    nbr[g] = point_leaf(t, px, py, pz);
}

hipError_t launch_neighbors(Tree t, int* nbr27, int nbricks, hipStream_t s) {
    if (nbricks <= 0) return hipSuccess;
    u64 n = u64(nbricks) * 27;
    hipLaunchKernelGGL(k_neighbors, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, t, nbr27, nbricks);
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

// One 512-thread workgroup per brick; threads 0..487 each own one apron voxel
// of the padded 10^3 brick (apron 1, 8^3 interior).  Enumeration: the z = 0
// and z = 9 faces (100 each), then the 36-voxel rings of z = 1..8.
template <typename AT>
__global__ __launch_bounds__(512) void k_apron(Tree t, const int* nbr, AT* atlas, float boundval) {
    const int b = blockIdx.x;
    const int a = threadIdx.x;
    const int br = t.brick_res;        // 10
    const int w = t.brickwid;          // 8
    const int face = br * br;
    const int ring = br * br - w * w;
    if (a >= 2 * face + w * ring) return;
    const Node* n = t.nodes[0] + b;
    if (!n->mFlags || n->mValue[0] < 0) return;
    int x, y, z;
    if (a < 2 * face) {
        z = a < face ? 0 : br - 1;
        int r = a % face;
        y = r / br;
        x = r - y * br;
    } else {
        int s = a - 2 * face;
        z = s / ring + 1;
        int r = s - (z - 1) * ring;
        if (r < br) { x = r; y = 0; }
        else if (r < 2 * br) { x = r - br; y = br - 1; }
        else { int q = r - 2 * br; y = q / 2 + 1; x = (q & 1) ? br - 1 : 0; }
    }
    int dx = x == 0 ? -1 : (x == br - 1 ? 1 : 0);
    int dy = y == 0 ? -1 : (y == br - 1 ? 1 : 0);
    int dz = z == 0 ? -1 : (z == br - 1 ? 1 : 0);
    int d = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
    int nb = nbr[u64(b) * 27 + d];
    u64 own = slot_of(t, n->mValue) * u64(t.brick_vox) + u64((z * br + y) * br + x);
    float v = boundval;
    if (nb >= 0) {
        const Node* m = t.nodes[0] + nb;
        // interior coordinate inside the neighbour, shifted by one for the apron
        int nx = (x - 1) - dx * w + 1, ny = (y - 1) - dy * w + 1, nz = (z - 1) - dz * w + 1;
        u64 src = slot_of(t, m->mValue) * u64(t.brick_vox) + u64((nz * br + ny) * br + nx);
        v = ld<AT>(atlas, src);
    }
    st<AT>(atlas, own, v);
}

hipError_t launch_update_apron(Tree t, const int* nbr27, int nbricks, void* atlas, int atlas_type, float boundval,
                               hipStream_t s) {
    if (nbricks <= 0) return hipSuccess;
    if (t.apron != 1) return hipErrorInvalidValue;
    if (atlas_type == T_HALF)
        hipLaunchKernelGGL(k_apron<__half>, dim3(nbricks), dim3(512), 0, s, t, nbr27, (__half*)atlas, boundval);
    else
        hipLaunchKernelGGL(k_apron<float>, dim3(nbricks), dim3(512), 0, s, t, nbr27, (float*)atlas, boundval);
    return hipGetLastError();
}

}  // namespace gvdb
