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

// A few small records travel inside the kernel arguments (no staging copy,
// no extra launch); larger sets go through a device table.
static const int MULTI_INLINE = 8;
static const int MULTI_INLINE_BYTES = 2048;

struct MultiMap {
    int G;
    const uint32_t* first;   // G + 1 entries (device), or null: inline
    const void* table;       // G argument records (device), or null: inline
    int xcd;                 // 1: workgroups take bricks in XCD-contiguous order (xcd_remap)
    uint32_t ifirst[MULTI_INLINE + 1];
    alignas(16) unsigned char itab[MULTI_INLINE_BYTES];
};

__device__ inline const uint32_t* multi_first(const MultiMap& m) { return m.first ? m.first : m.ifirst; }

template <class A>
__device__ inline const A& multi_rec(const MultiMap& m, int g) {
    return reinterpret_cast<const A*>(m.table ? m.table : static_cast<const void*>(m.itab))[g];
}

// Launches set up with inline_ok = false keep first[] and the records in
// device memory; kernels that index them only through these pointers never
// address the inline arrays of the by-value map (which a runtime index can
// make the compiler copy to scratch, per lane).
template <class A>
__device__ inline const A& multi_rec_dev(const MultiMap& m, int g) {
    return static_cast<const A*>(m.table)[g];
}
// The grid of workgroup b: the last g with first[g] <= b (first[] ascending,
// first[0] = 0), counted 64 entries at a time, one load per lane and a
// ballot, instead of a binary search's chain of dependent loads (which a
// workgroup with little work pays in full).  Called by whole wave64s, every
// lane active.
__device__ inline int multi_grid_in(const uint32_t* first, int G, uint32_t b) {
    const int lane = int(threadIdx.x & 63u);
    int g = 0;
    for (int base = 1; base < G; base += 64) {
        const int i = base + lane;
        const uint64_t bal = __ballot(i < G && first[i] <= b);
        g += __popcll(bal);
        if (bal != ~0ull) break;
    }
    return g;
}
__device__ inline int multi_grid_dev(const MultiMap& m, uint32_t b) { return multi_grid_in(m.first, m.G, b); }

__device__ inline int multi_grid(const MultiMap& m, uint32_t b) { return multi_grid_in(multi_first(m), m.G, b); }

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
        const uint32_t bi = m.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;               \
        if (m.G <= 1) {                                                                          \
            NAME##_body(a0, bi, gridDim.x);                                                      \
            return;                                                                              \
        }                                                                                        \
        const int g = multi_grid(m, bi);                                                         \
        const uint32_t* first = multi_first(m);                                                  \
        NAME##_body(multi_rec<A>(m, g), bi - first[g], first[g + 1] - first[g]);                 \
    }

// Same for a body templated on one parameter (TDECL e.g. `typename AT`).
#define GVDB_MULTI_T(NAME, TDECL, TNAME, A, THREADS)                                             \
    template <TDECL>                                                                             \
    __global__ __launch_bounds__(THREADS) void NAME(A a0, MultiMap m) {                          \
        const uint32_t bi = m.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;               \
        if (m.G <= 1) {                                                                          \
            NAME##_body<TNAME>(a0, bi, gridDim.x);                                               \
            return;                                                                              \
        }                                                                                        \
        const int g = multi_grid(m, bi);                                                         \
        const uint32_t* first = multi_first(m);                                                  \
        NAME##_body<TNAME>(multi_rec<A>(m, g), bi - first[g], first[g + 1] - first[g]);          \
    }

// Pinned staging for argument tables: bump-allocated, copied to the device
// on the launch stream before the kernel that reads them.  Reset only after
// the stream has drained (the owner knows when).
struct Arena {
    uint8_t* host = nullptr;
    uint8_t* dev = nullptr;
    size_t cap = 0, used = 0;
    // a captured graph's arena: never wraps (wrapping waits for the stream),
    // so its tables stay as the graph's copies read them at every replay
    bool fixed = false;
    hipError_t init(size_t bytes);
    void release();
    // Returns the device address of a copy of src[0..n), or null on failure.
    void* push(const void* src, size_t n, hipStream_t s, hipError_t& e);
    void reset() { used = 0; }
};

// Launch geometry for G grids needing nb[g] blocks each.
struct MultiLaunch {
    MultiMap map{1, nullptr, nullptr, 0, {}, {}};
    uint32_t blocks = 0;
};

hipError_t multi_setup_raw(Arena* ar, const void* args, size_t rec, const uint32_t* nb, int G, hipStream_t s,
                           MultiLaunch& out,
                           bool inline_ok = true);

template <typename A>
inline hipError_t multi_setup(Arena* ar, const A* args, const uint32_t* nb, int G, hipStream_t s, MultiLaunch& out,
                              bool inline_ok = true) {
    return multi_setup_raw(ar, args, sizeof(A), nb, G, s, out, inline_ok);
}

}  // namespace gvdb
