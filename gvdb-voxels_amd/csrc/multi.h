// multi.h -- one kernel launch over many grids.
//
// Every kernel of the path has an argument record A and a body
// `NAME_body(const A&, uint32_t block, uint32_t nblocks)` written for one
// grid.  GVDB_MULTI(NAME, A, THREADS) defines the kernel NAME(A a0, MultiMap m):
//   m.G <= 1: one grid, arguments by value, block = blockIdx.x;
//   m.G  > 1: G grids in one launch; grid g owns blocks [first[g], first[g+1])
//             and its record is table[g] (device memory), so small and large
//             grids share one launch and fill the device together.
// The body is instantiated once per case so the single-grid path keeps its
// arguments in the kernel-argument segment.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "gvdb_device.h"

namespace gvdb {

struct MultiMap {
    int G;
    const uint32_t* first;   // G + 1 entries
    const void* table;       // G argument records
};

__device__ inline int multi_grid(const MultiMap& m, uint32_t b) {
    int lo = 0, hi = m.G - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (m.first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Workgroups are handed to the 8 XCDs of the device in turn; this maps the
// b-th workgroup of a launch to the brick order so that each XCD (with its
// own L2) works through one contiguous eighth of the bricks, where the
// neighbours of a brick (nearby in pool order) share its cache.
__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t n) {
    const uint32_t q = n >> 3;
    if (b >= (q << 3)) return b;
    return (b & 7u) * q + (b >> 3);
}

#define GVDB_MULTI(NAME, A, THREADS)                                                             \
    __global__ __launch_bounds__(THREADS) void NAME(A a0, MultiMap m) {                          \
        if (m.G <= 1) {                                                                          \
            NAME##_body(a0, blockIdx.x, gridDim.x);                                              \
            return;                                                                              \
        }                                                                                        \
        const int g = multi_grid(m, blockIdx.x);                                                 \
        const A* tab = reinterpret_cast<const A*>(m.table);                                      \
        NAME##_body(tab[g], blockIdx.x - m.first[g], m.first[g + 1] - m.first[g]);               \
    }

// Same for a body templated on one parameter (TDECL e.g. `typename AT`).
#define GVDB_MULTI_T(NAME, TDECL, TNAME, A, THREADS)                                             \
    template <TDECL>                                                                             \
    __global__ __launch_bounds__(THREADS) void NAME(A a0, MultiMap m) {                          \
        if (m.G <= 1) {                                                                          \
            NAME##_body<TNAME>(a0, blockIdx.x, gridDim.x);                                       \
            return;                                                                              \
        }                                                                                        \
        const int g = multi_grid(m, blockIdx.x);                                                 \
        const A* tab = reinterpret_cast<const A*>(m.table);                                      \
        NAME##_body<TNAME>(tab[g], blockIdx.x - m.first[g], m.first[g + 1] - m.first[g]);        \
    }

// Pinned staging for argument tables: bump-allocated, copied to the device
// on the launch stream before the kernel that reads them.  Reset only after
// the stream has drained (the owner knows when).
struct Arena {
    uint8_t* host = nullptr;
    uint8_t* dev = nullptr;
    size_t cap = 0, used = 0;
    hipError_t init(size_t bytes);
    void release();
    // Returns the device address of a copy of src[0..n), or null on failure.
    void* push(const void* src, size_t n, hipStream_t s, hipError_t& e);
    void reset() { used = 0; }
};

// Launch geometry for G grids needing nb[g] blocks each.
struct MultiLaunch {
    MultiMap map{1, nullptr, nullptr};
    uint32_t blocks = 0;
};

hipError_t multi_setup_raw(Arena* ar, const void* args, size_t rec, const uint32_t* nb, int G, hipStream_t s,
                           MultiLaunch& out);

template <typename A>
inline hipError_t multi_setup(Arena* ar, const A* args, const uint32_t* nb, int G, hipStream_t s, MultiLaunch& out) {
    return multi_setup_raw(ar, args, sizeof(A), nb, G, s, out);
}

}  // namespace gvdb
