// gvdb_oracle.cpp -- CPU restatement of GVDB's point-to-sparse-volume path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke()
// entry point and bench.py's cpu_baseline leg may load it.  It is never part
// of the product path.
//
// Every routine below restates the reference algorithm (NVIDIA GVDB Voxels
// 1.1.1) from its source, in plain host C++ with no GPU; each one cites the
// reference file:line it follows.  Paths are relative to
// source/gvdb_library/ of the reference.
//
// Two arithmetic variants exist for the numerical stages:
//   * FP32 -- the reference's own float arithmetic (what its kernels compute),
//             used as the timed CPU baseline;
//   * FP64 -- the same loops in double precision, used as the error reference.
// Topology, counts and indices are exact integers in both.
//
// Where the reference reads memory out of bounds (child-list slots outside
// [0,res) in its tree walks), the oracle does not guess: it records the event
// in `undefined_events` so a test can reject that input as outside the
// reference's defined behaviour.
//
// Build: see oracle/Makefile (g++ -O2 -fopenmp, shared library with a C ABI).

#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>
#include <numeric>
#include <atomic>
#include <chrono>
#include <parallel/algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned long long u64;
typedef long long i64;

static const u64 ID_UNDEFL = 0xFFFFFFFFull;           // src/gvdb_types.h, kernels/cuda_gvdb_nodes.cuh:19
static const u64 ID_UNDEF64 = 0xFFFFFFFFFFFFFFFFull;  // kernels/cuda_gvdb_nodes.cuh:20
static const float NOHIT = 1.0e10f;                   // kernels/cuda_gvdb_scene.cuh:20
static const int MAXLEV_HOST = 10;                    // src/gvdb_volume_gvdb.h:40
static const int MAX_ITER = 256;                      // kernels/cuda_gvdb_raycast.cuh:24
static const int TF_SIZE = 16384;                     // src/gvdb_scene.cpp:68

// Pool reference packing: src/gvdb_allocator.h:59-62
static inline u64 Elem(int grp, int lev, u64 ndx) { return u64(grp) | (u64(lev) << 8) | (ndx << 16); }
static inline int ElemLev(u64 id) { return int((id >> 8) & 0xFF); }
static inline u64 ElemNdx(u64 id) { return id >> 16; }

// 64-byte node: src/gvdb_node.h:27-40
struct Node {
    uint8_t mLev, mFlags, mPriority, pad;
    int32_t mPos[3];
    int32_t mValue[3];
    float mVRange[3];
    u64 mParent;
    u64 mChildList;
    u64 mMask;
};
static_assert(sizeof(Node) == 64, "node is 64 bytes");

// Atlas map entry: kernels/cuda_gvdb_nodes.cuh:38-41 (AtlasNode in gvdb_volume_gvdb.h)
struct AtlasNode { int32_t mPos[3]; int32_t mLeafNode; };

// One pool: src/gvdb_allocator.cpp:55-91 (PoolCreate), :163-192 (PoolAlloc)
struct Pool {
    bool exists = false;
    u64 stride = 0, max = 0, lastEle = 0, usedNum = 0;
    std::vector<uint8_t> data;
};

struct Vol {
    int levels = 0;
    int logdim[MAXLEV_HOST] = {0};
    int res[MAXLEV_HOST] = {0};
    i64 range[MAXLEV_HOST] = {0};
    Pool pool[2][MAXLEV_HOST];
    u64 root = ID_UNDEFL;
    bool rebuild_topo = true;
    int curr_depth = -1;
    float posMin[3] = {0, 0, 0}, posMax[3] = {0, 0, 0};
    float objMin[3] = {0, 0, 0}, objMax[3] = {0, 0, 0};
    int top_lev = 1;
    float epsilon = 0.001f;
    std::atomic<int> undefined_events{0};
    // atlas (one float channel; the FP64 shadow copy serves the FP64 oracle)
    int default_axiscnt[3] = {16, 16, 1};
    bool has_atlas = false;
    int apron = 1;
    int brickwid = 8;
    int atlas_cnt[3] = {0, 0, 0};
    u64 atlas_max = 0, atlas_last = 0, atlas_used = 0;
    std::vector<float> atlasF;
    std::vector<double> atlasD;
    // one uchar4 channel (the colour channel of GatherDensity's chanClr and
    // of SetColorChannel), same slot layout, a uint32 per voxel
    bool has_clr = false;
    std::vector<uint32_t> atlasC;
    bool render_clr = false;   // SetColorChannel on that channel
    bool fp64 = true;      // keep the FP64 shadow atlas (off for the timed CPU baseline)
    std::vector<AtlasNode> amap;
    u64 amap_max = 0;
    // InsertPointsSubcell outputs
    std::vector<int> sc_cnt, sc_off, sc_nid, sc_mapping, sc_pos;
    std::vector<float> sc_pnt;
    std::vector<uint32_t> sc_clr;   // per-entry colours (SetPoints' colour buffer), or empty
    int sc_len = 0;
    // InsertPointsSubcell_FP16 entries and the bounds they are quantised against
    std::vector<uint16_t> sc_pnt16;
    bool sc16 = false;     // sc_pnt16 belongs to the current entry lists
    float qmin[3] = {0, 0, 0}, qrange[3] = {0, 0, 0};
    int num_sc = 0;
    // timing of the last call of each stage (ms), for the CPU baseline
    double stage_ms[8] = {0};
};

// ---------------------------------------------------------------- pools ----
static void PoolCreate(Vol& v, int grp, int lev, u64 width, u64 initmax) {
    // src/gvdb_allocator.cpp:55-91: a zero-size pool is a placeholder
    Pool p;
    p.stride = width;
    p.max = initmax;
    if (width * initmax == 0) { v.pool[grp][lev] = p; return; }
    p.exists = true;
    p.data.assign(width * initmax, 0);
    v.pool[grp][lev] = p;
}

static u64 PoolAlloc(Vol& v, int grp, int lev) {
    // src/gvdb_allocator.cpp:163-192: doubling growth, returns Elem(grp,lev,idx)
    if (lev >= v.levels) return ID_UNDEFL;
    Pool& p = v.pool[grp][lev];
    if (!p.exists) return ID_UNDEFL;
    if (p.lastEle >= p.max) {
        p.max *= 2;
        p.data.resize(p.stride * p.max, 0);
    }
    p.lastEle++;
    p.usedNum++;
    return Elem(grp, lev, p.lastEle - 1);
}

static inline Node* nodeAt(Vol& v, int lev, u64 idx) {
    return reinterpret_cast<Node*>(&v.pool[0][lev].data[idx * 64]);
}
static inline Node* getNode(Vol& v, u64 ref) { return nodeAt(v, ElemLev(ref), ElemNdx(ref)); }
static inline u64* PoolData64(Vol& v, u64 ref) {
    int grp = int(ref & 0xFF), lev = ElemLev(ref);
    Pool& p = v.pool[grp][lev];
    return reinterpret_cast<u64*>(&p.data[ElemNdx(ref) * p.stride]);
}

static void SetupNode(Vol& v, u64 nodeid, int lev, const int pos[3], bool marker = true) {
    // src/gvdb_volume_gvdb.cpp:2512-2524
    Node* n = getNode(v, nodeid);
    n->mLev = uint8_t(lev);
    n->mPos[0] = pos[0]; n->mPos[1] = pos[1]; n->mPos[2] = pos[2];
    n->mChildList = ID_UNDEFL;
    n->mParent = ID_UNDEFL;
    n->mValue[0] = n->mValue[1] = n->mValue[2] = -1;
    n->mFlags = marker ? 1 : 0;
}

static u64 voxcnt(const Vol& v, int lev) { u64 r = u64(1) << v.logdim[lev]; return r * r * r; }

static void Clear(Vol& v) {
    // src/gvdb_volume_gvdb.cpp:2487-2503, src/gvdb_allocator.cpp:194-203,683-690
    for (int g = 0; g < 2; g++)
        for (int l = 0; l < v.levels; l++) { v.pool[g][l].usedNum = 0; v.pool[g][l].lastEle = 0; }
    v.root = ID_UNDEFL;
    v.atlas_last = v.atlas_used = 0;
    v.rebuild_topo = true;
}

static void Configure(Vol& v, int levs, const int* r, const int* numcnt) {
    // src/gvdb_volume_gvdb.cpp:2380-2434 (N-level form)
    v.levels = levs;
    for (int n = 0; n < levs; n++) {
        v.logdim[n] = (r[n] == 0) ? 1 : r[n];
        v.res[n] = 1 << v.logdim[n];
    }
    // getRange: src/gvdb_volume_gvdb.h:614-619
    for (int l = 0; l < levs; l++) {
        i64 rr = v.res[0];
        for (int k = 1; k <= l; k++) rr *= v.res[k];
        v.range[l] = rr;
    }
    for (int n = 0; n < levs; n++) PoolCreate(v, 0, n, 64, numcnt[n] == 0 ? 1 : numcnt[n]);
    PoolCreate(v, 1, 0, 0, 0);
    for (int n = 1; n < levs; n++) PoolCreate(v, 1, n, sizeof(u64) * voxcnt(v, n), numcnt[n] == 0 ? 1 : numcnt[n]);
    Clear(v);
}

// Host GetCoveringNode with integer position: src/gvdb_volume_gvdb.cpp:2915-2931
static void GetCoveringNodeHost(const Vol& v, int lev, const int pos[3], int out[3]) {
    i64 rg = v.range[lev];
    if (lev == MAXLEV_HOST - 1) { out[0] = out[1] = out[2] = 0; return; }
    for (int a = 0; a < 3; a++) {
        i64 np = (i64(pos[a]) / rg) * rg;     // C++ truncating division, as Vector3DI /=
        if (pos[a] < np) np -= rg;
        out[a] = int(np);
    }
}

// Device GetCoveringNode(float3, int): kernels/cuda_gvdb_particles.cuh:88-100
static inline int coverF(float p, int range) {
    int np = int(ceilf(p / float(range)) * float(range));
    if (p < float(np)) np -= range;
    return np;
}

// -------------------------------------------------- device-style tree walks
// getChild: kernels/cuda_gvdb_nodes.cuh:115-124 (dense child lists).  Returns
// the child index as the reference's `int` (-1 for an empty slot).
static inline int getChildD(Vol& v, Node* node, int b) {
    u64 listid = node->mChildList;
    if (listid == ID_UNDEFL) return -1;
    int clev = int((listid >> 8) & 0xFF);
    u64 cndx = listid >> 16;
    Pool& p = v.pool[1][clev];
    u64* clist = reinterpret_cast<u64*>(&p.data[cndx * p.stride]);
    return int(clist[b] >> 16);
}

static inline bool inRes(Vol& v, const int pin[3], int res) {
    bool ok = pin[0] >= 0 && pin[1] >= 0 && pin[2] >= 0 && pin[0] < res && pin[1] < res && pin[2] < res;
    if (!ok) v.undefined_events++;
    return ok;
}

// getParent: kernels/cuda_gvdb_nodes.cuh:350-397.  Returns the parent index,
// or -2 where the reference's own code would walk through an empty slot
// (its ID_UNDEFL test compares a sign-extended -1 with 0xFFFFFFFF and misses).
static i64 getParentD(Vol& v, int stopLev, const int pos[3]) {
    Node* nd = nodeAt(v, v.top_lev, 0);
    if (!nd->mFlags) return 0;
    int l = nd->mLev;
    if (l == stopLev) return 0;
    while (l > stopLev) {
        int res = v.res[l];
        i64 rg = v.range[l];
        int pin[3];
        for (int a = 0; a < 3; a++) pin[a] = int(((i64)(pos[a] - nd->mPos[a]) * res) / rg);
        if (!inRes(v, pin, res)) return -2;
        int bit = (pin[2] * res + pin[1]) * res + pin[0];
        l--;
        int child = getChildD(v, nd, bit);
        if (child < 0) { v.undefined_events++; return -2; }
        if (l == stopLev) return child;
        nd = nodeAt(v, l, u64(child));
    }
    return -2;
}

// getPosParent: kernels/cuda_gvdb_nodes.cuh:303-348 (returns -1 when missing)
static i64 getPosParentD(Vol& v, const int pos[3], int stopLev) {
    Node* nd = nodeAt(v, v.top_lev, 0);
    int l = nd->mLev;
    if (l == 0) return 0;
    while (l > 0) {
        int res = v.res[l];
        i64 rg = v.range[l];
        int pin[3];
        for (int a = 0; a < 3; a++) pin[a] = int(((i64)(pos[a] - nd->mPos[a]) * res) / rg);
        if (!inRes(v, pin, res)) return -1;
        int bit = (pin[2] * res + pin[1]) * res + pin[0];
        int child = getChildD(v, nd, bit);
        l--;
        if (child < 0) return -1;
        if (l == stopLev) return child;
        nd = nodeAt(v, l, u64(child));
    }
    return -1;
}

// getPosLeafParent: kernels/cuda_gvdb_nodes.cuh:255-288 (returns -1 when missing)
static int getPosLeafParentD(Vol& v, const int pos[3]) {
    Node* nd = nodeAt(v, v.top_lev, 0);
    int l = nd->mLev;
    if (l == 0) return 0;
    while (true) {
        int res = v.res[l];
        i64 rg = v.range[l];
        int pin[3];
        for (int a = 0; a < 3; a++) pin[a] = int(((i64)(pos[a] - nd->mPos[a]) * res) / rg);
        // the reference tests only `> res`; anything outside [0,res) reads
        // outside the child list, which the oracle reports and treats as absent
        if (!inRes(v, pin, res)) return -1;
        int bit = (pin[2] * res + pin[1]) * res + pin[0];
        int child = getChildD(v, nd, bit);
        if (child < 0) return -1;
        l--;
        if (l == 0) return child;
        nd = nodeAt(v, l, u64(child));
    }
}

// getNode(gvdb, lev, start_id, pos, &id) walk with float bounds:
// kernels/cuda_gvdb_nodes.cuh:199-226, used by getNodeAtPoint (:243-253).
template <typename T>
static i64 getNodeAtPointT(Vol& v, const T pos[3]) {
    int lev = v.top_lev;
    Node* node = nodeAt(v, lev, 0);
    i64 id = 0;
    while (lev > 0) {
        T vmin[3] = {T(node->mPos[0]), T(node->mPos[1]), T(node->mPos[2])};
        T rg = T(v.range[lev]);
        for (int a = 0; a < 3; a++)
            if (pos[a] < vmin[a] || pos[a] >= vmin[a] + rg) return -1;
        T vdel = T(v.range[lev] / v.res[lev]);
        int p[3];
        for (int a = 0; a < 3; a++) p[a] = int((pos[a] - vmin[a]) / vdel);
        int b = ((p[2] << v.logdim[lev]) + p[1]) << v.logdim[lev];
        b += p[0];
        lev--;
        int child = getChildD(v, node, b);
        if (child < 0) return -1;
        id = child;
        node = nodeAt(v, lev, u64(child));
    }
    return id;
}

// PrepareVDBPartially / PrepareVDB top level: src/gvdb_volume_gvdb.cpp:3955-3979, :3999-4013
static void computeTopLev(Vol& v) {
    int tlev = 1;
    for (int n = v.levels - 1; n >= 0; n--)
        if (v.pool[0][n].lastEle == 1) tlev = n;
    v.top_lev = tlev;
}

// --------------------------------------------------------- GPU topology ----
static u64 makeKey(int lev, int x, int y, int z) {
    // kernels/cuda_gvdb_particles.cuh:377-380: u16 {z, y, x, lev} read as one u64
    return (u64(uint16_t(lev)) << 48) | (u64(uint16_t(x)) << 32) | (u64(uint16_t(y)) << 16) | u64(uint16_t(z));
}

// FindUniqueBrick: src/gvdb_volume_gvdb.cpp:843-873, kernels/cuda_gvdb_particles.cuh:336-350,934-943
static void findUnique(std::vector<u64>& keys, int depth, std::vector<u64>& uniq, std::vector<int>& levCnt) {
    // gvdbDeviceRadixSort: ascending u64 (src/gvdb_cutils.cu:11-16); any
    // correct sort gives the same sequence, so the oracle sorts in parallel
    __gnu_parallel::sort(keys.begin(), keys.end());
    uniq.clear();
    levCnt.assign(depth, 0);
    for (size_t i = 0; i < keys.size(); i++) {
        if (i == 0 || keys[i] != keys[i - 1]) {
            uniq.push_back(keys[i]);
            int lev = int((keys[i] >> 48) & 0xFF);
            if (lev != 255) levCnt[lev]++;
        }
    }
}

// Node creation from the unique list: src/gvdb_volume_gvdb.cpp:1308-1331 and :965-988
static void createNodes(Vol& v, int depth, const std::vector<u64>& uniq, const std::vector<int>& levCnt) {
    int pre = 0;
    for (int lev = 0; lev < depth; lev++) {
        i64 rg = v.range[lev];
        for (int n = pre; n < levCnt[lev] + pre; n++) {
            u64 key = uniq[n];
            int pos[3] = {int(((key >> 32) & 0xFFFF) * rg), int(((key >> 16) & 0xFFFF) * rg), int((key & 0xFFFF) * rg)};
            u64 id = PoolAlloc(v, 0, lev);
            SetupNode(v, id, lev, pos);
            if (lev > 0) {
                Node* nd = getNode(v, id);
                nd->mChildList = PoolAlloc(v, 1, lev);
                u64* cl = PoolData64(v, nd->mChildList);
                memset(cl, 0xFF, sizeof(u64) * voxcnt(v, lev));
            }
        }
        pre += levCnt[lev];
    }
}

// gvdbLinkBricks: kernels/cuda_gvdb_nodes.cuh:470-505, launched per level
// top-down at src/gvdb_volume_gvdb.cpp:1395-1401 and :999-1006
static void linkLevels(Vol& v, int depth) {
    computeTopLev(v);
    for (int lev = depth - 1; lev >= 0; lev--) {
        u64 cnt = v.pool[0][lev].lastEle;
        u64 pcnt = v.pool[0][lev + 1].lastEle;
        for (u64 i = 0; i < cnt; i++) {
            Node* node = nodeAt(v, lev, i);
            if (!node->mFlags) continue;
            i64 pid = getParentD(v, lev + 1, node->mPos);
            if (pid < 0) { v.undefined_events++; continue; }
            Node* pnode = nodeAt(v, lev + 1, u64(pid));
            int res = v.res[lev + 1];
            i64 rg = v.range[lev + 1];
            int pin[3];
            for (int a = 0; a < 3; a++) pin[a] = int(((i64)(node->mPos[a] - pnode->mPos[a]) * res) / rg);
            int bit = (pin[2] * res + pin[1]) * res + pin[0];
            if (pin[0] > res || pin[0] < 0 || pin[1] > res || pin[1] < 0 || pin[2] > res || pin[2] < 0) continue;
            node->mParent = (u64(pid) << 16) | (u64(lev + 1) << 8);
            u64 cndx = pnode->mChildList >> 16;
            if (cndx >= pcnt) continue;
            Pool& cp = v.pool[1][lev + 1];
            u64* cl = reinterpret_cast<u64*>(&cp.data[cndx * cp.stride]);
            cl[bit] = (u64(i) << 16) | (u64(lev) << 8);
        }
    }
}

// GetBoundingBox: src/gvdb_volume_gvdb.cpp:5404-5467 (min/max of pos+trans, padded)
static void boundingBox(Vol& v, const float* pts, int n, const float orig[3]) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            float w = pts[3 * i + a] + orig[a];
            mn[a] = std::min(mn[a], w);
            mx[a] = std::max(mx[a], w);
        }
    for (int a = 0; a < 3; a++) {
        v.posMin[a] = mn[a] - 1;
        v.posMax[a] = mx[a] + 2;
        v.posMin[a] = 0 > v.posMin[a] ? 0 : v.posMin[a];
        v.posMax[a] = 0 > v.posMax[a] ? 0 : v.posMax[a];
    }
}

// DetermineDepth: src/gvdb_volume_gvdb.cpp:821-837
static int determineDepth(Vol& v, int rootPos[3]) {
    int pmin[3] = {int(v.posMin[0]), int(v.posMin[1]), int(v.posMin[2])};
    int pmax[3] = {int(v.posMax[0]), int(v.posMax[1]), int(v.posMax[2])};
    for (int lev = 0; lev < v.levels; lev++) {
        int a[3], b[3];
        GetCoveringNodeHost(v, lev, pmin, a);
        GetCoveringNodeHost(v, lev, pmax, b);
        if (a[0] == b[0] && a[1] == b[1] && a[2] == b[2]) {
            rootPos[0] = a[0]; rootPos[1] = a[1]; rootPos[2] = a[2];
            return lev;
        }
    }
    return -1;
}

static void activateIncre(Vol& v, const float* pts, int n, float radius, const float orig[3], int D, bool accum);

// RebuildTopology: src/gvdb_volume_gvdb.cpp:1181-1218.  A full rebuild
// (ActivateBricksGPU :1229-1413, ActivateExtraBricksGPU :931-1016) after
// RequestFullRebuild or Configure, or when the depth differs from mCurrDepth
// (which only that branch sets, so the second call after Configure rebuilds
// fully too); otherwise the incremental ActivateIncreBricksGPU without
// accumulation.  An incremental call whose depth is not the root's level
// (a full rebuild at another depth since mCurrDepth was set) rebuilds fully
// here; the reference would walk a tree of another depth.
// Returns 0 on success, a negative code on the reference's error exits.
static int rebuildTopology(Vol& v, const float* pts, int n, float radius, const float orig[3]) {
    boundingBox(v, pts, n, orig);
    int rootPos[3];
    int depth = determineDepth(v, rootPos);
    if (depth <= 0) return -4;                  // :1187-1190 (and "not found")
    if (depth > 5 || depth >= v.levels) return -4;  // :1191-1194, and no pool above the top level
    v.undefined_events = 0;
    if (v.rebuild_topo) {
        Clear(v);
    } else if (depth != v.curr_depth || v.root == ID_UNDEFL || ElemLev(v.root) != depth) {
        Clear(v);
        v.rebuild_topo = true;
        v.curr_depth = depth;
    }
    if (!v.rebuild_topo) {
        activateIncre(v, pts, n, radius, orig, depth, false);
        return 0;
    }

    // gvdbCalcBrickId: kernels/cuda_gvdb_particles.cuh:352-382
    std::vector<u64> keys((size_t)n * depth);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        float w[3] = {pts[3 * i] + orig[0], pts[3 * i + 1] + orig[1], pts[3 * i + 2] + orig[2]};
        bool neg = w[0] < 0 || w[1] < 0 || w[2] < 0;
        for (int lev = 0; lev < depth; lev++) {
            if (neg) { keys[(size_t)i * depth + lev] = makeKey(255, 0, 0, 0); continue; }
            int rr = int(v.range[lev]);
            int bx = coverF(w[0], rr), by = coverF(w[1], rr), bz = coverF(w[2], rr);
            keys[(size_t)i * depth + lev] = makeKey(lev, bx / rr, by / rr, bz / rr);
        }
    }
    std::vector<u64> uniq;
    std::vector<int> levCnt;
    findUnique(keys, depth, uniq, levCnt);

    // root node and its child list: :1289-1300
    v.root = PoolAlloc(v, 0, depth);
    u64 cl = PoolAlloc(v, 1, depth);
    SetupNode(v, v.root, depth, rootPos);
    Node* rn = getNode(v, v.root);
    rn->mFlags = 1;
    rn->mChildList = cl;
    memset(PoolData64(v, cl), 0xFF, sizeof(u64) * voxcnt(v, depth));
    createNodes(v, depth, uniq, levCnt);
    linkLevels(v, depth);

    if (radius >= 1) {
        // gvdbCalcExtraBrickId: kernels/cuda_gvdb_particles.cuh:467-512
        // the lookups only read the tree, and the keys are sorted and made
        // unique afterwards, so points are split over threads
        std::vector<u64> ex;
#pragma omp parallel
        {
        std::vector<u64> mine;
#pragma omp for schedule(static) nowait
        for (int i = 0; i < n; i++) {
            float w[3] = {pts[3 * i] + orig[0], pts[3 * i + 1] + orig[1], pts[3 * i + 2] + orig[2]};
            if (w[0] < 0 || w[1] < 0 || w[2] < 0) continue;
            int bmin[3], bmax[3];
            for (int a = 0; a < 3; a++) {
                bmin[a] = int(float(int(w[a])) - radius);
                bmax[a] = int(float(int(w[a])) + 2 * radius);
            }
            for (int lev = 0; lev < depth; lev++) {
                int rr = int(v.range[lev]);
                int nd[3];
                for (nd[0] = bmin[0] / rr * rr; nd[0] <= bmax[0] / rr * rr; nd[0] += rr)
                    for (nd[1] = bmin[1] / rr * rr; nd[1] <= bmax[1] / rr * rr; nd[1] += rr)
                        for (nd[2] = bmin[2] / rr * rr; nd[2] <= bmax[2] / rr * rr; nd[2] += rr) {
                            if (getPosParentD(v, nd, lev) < 0)
                                mine.push_back(makeKey(lev, nd[0] / rr, nd[1] / rr, nd[2] / rr));
                        }
            }
        }
        // drop repeats before merging (the set is what counts)
        std::sort(mine.begin(), mine.end());
        mine.erase(std::unique(mine.begin(), mine.end()), mine.end());
#pragma omp critical
        ex.insert(ex.end(), mine.begin(), mine.end());
        }
        std::vector<u64> xu;
        std::vector<int> xc;
        findUnique(ex, depth, xu, xc);
        createNodes(v, depth, xu, xc);
        linkLevels(v, depth);
    }
    v.rebuild_topo = false;
    computeTopLev(v);
    return 0;
}

// ----------------------------------------------------------- host topology
// ActivateIncreBricksGPU: src/gvdb_volume_gvdb.cpp:1020-1156.  Leaf markers
// (the leaves' flags when accumulating, else 0); every point w = p + orig >= 0
// looks up, at each level below D, the nodes covering [int(w) - r, int(w) + r]
// (gvdbCalcIncreExtraBrickId, kernels/cuda_gvdb_particles.cuh:420-465):
// missing ones become keys, found leaves get marker 1.  With no key nothing
// changes (the markers are not applied either).  Otherwise the leaves take
// their markers, the unique keys become nodes appended in key order, every
// level is linked top-down (gvdbLinkBricks), unflagged leaves are delinked
// (gvdbDelinkLeafBricks, kernels/cuda_gvdb_nodes.cuh:401-429) and then each
// level 1..D-1 bottom-up (gvdbDelinkBricks, :431-467), and the leaf pool's
// used count drops by the unflagged leaves.  Node positions below zero
// (radius 8 or more) would go through the reference's signed-to-u16 key
// cast; they are counted as undefined and skipped.
static void activateIncre(Vol& v, const float* pts, int n, float radius, const float orig[3], int D, bool accum) {
    const u64 total = v.pool[0][0].lastEle;
    if (total == 0) return;
    std::vector<int> marker(total, 0);
    if (accum)
        for (u64 i = 0; i < total; i++) marker[i] = nodeAt(v, 0, i)->mFlags;
    std::vector<u64> ex;
    for (int i = 0; i < n; i++) {
        float w[3] = {pts[3 * i] + orig[0], pts[3 * i + 1] + orig[1], pts[3 * i + 2] + orig[2]};
        if (w[0] < 0 || w[1] < 0 || w[2] < 0) continue;
        int bmin[3], bmax[3];
        for (int a = 0; a < 3; a++) {
            bmin[a] = int(float(int(w[a])) - radius);
            bmax[a] = int(float(int(w[a])) + radius);
        }
        for (int lev = 0; lev < D; lev++) {
            int rr = int(v.range[lev]);
            int nd[3];
            for (nd[0] = bmin[0] / rr * rr; nd[0] <= bmax[0] / rr * rr; nd[0] += rr)
                for (nd[1] = bmin[1] / rr * rr; nd[1] <= bmax[1] / rr * rr; nd[1] += rr)
                    for (nd[2] = bmin[2] / rr * rr; nd[2] <= bmax[2] / rr * rr; nd[2] += rr) {
                        if (nd[0] < 0 || nd[1] < 0 || nd[2] < 0) { v.undefined_events++; continue; }
                        i64 id = getPosParentD(v, nd, lev);
                        if (id < 0) ex.push_back(makeKey(lev, nd[0] / rr, nd[1] / rr, nd[2] / rr));
                        else if (lev == 0) marker[size_t(id)] = 1;
                    }
        }
    }
    if (ex.empty()) return;
    std::vector<u64> uniq;
    std::vector<int> levCnt;
    findUnique(ex, D, uniq, levCnt);
    for (u64 i = 0; i < total; i++) nodeAt(v, 0, i)->mFlags = uint8_t(marker[i]);
    createNodes(v, D, uniq, levCnt);
    linkLevels(v, D);
    // gvdbDelinkLeafBricks
    auto clearSlot = [&](Node* node, int plev) {
        i64 pid = getParentD(v, plev, node->mPos);
        if (pid < 0) { v.undefined_events++; return; }
        Node* pn = nodeAt(v, plev, u64(pid));
        int res = v.res[plev];
        i64 rg = v.range[plev];
        int pin[3];
        for (int a = 0; a < 3; a++) pin[a] = int(((i64)(node->mPos[a] - pn->mPos[a]) * res) / rg);
        if (!inRes(v, pin, res)) return;
        node->mParent = ID_UNDEF64;
        u64* cl = PoolData64(v, pn->mChildList);
        cl[(pin[2] * res + pin[1]) * res + pin[0]] = ID_UNDEF64;
    };
    for (u64 i = 0; i < v.pool[0][0].lastEle; i++) {
        Node* node = nodeAt(v, 0, i);
        if (node->mFlags || node->mParent == ID_UNDEF64) continue;
        clearSlot(node, 1);
    }
    // gvdbDelinkBricks, levels 1 .. D-1
    for (int lev = 1; lev < D; lev++) {
        u64 cnt = v.pool[0][lev].lastEle;
        for (u64 i = 0; i < cnt; i++) {
            Node* node = nodeAt(v, lev, i);
            if (node->mParent == ID_UNDEF64) { node->mFlags = 0; continue; }
            u64* cl = reinterpret_cast<u64*>(&v.pool[1][lev].data[i * v.pool[1][lev].stride]);
            bool any = false;
            for (u64 c = 0; c < voxcnt(v, lev) && !any; c++) any = cl[c] != ID_UNDEF64;
            if (any) { node->mFlags = 1; continue; }
            node->mFlags = 0;
            clearSlot(node, lev + 1);
        }
    }
    // UpdateUsed: :1145-1149
    u64 used = v.pool[0][0].lastEle;
    for (u64 i = 0; i < v.pool[0][0].lastEle; i++)
        if (!nodeAt(v, 0, i)->mFlags) used--;
    v.pool[0][0].usedNum = used;
    computeTopLev(v);
}

// AccumulateTopology: src/gvdb_volume_gvdb.cpp:1158-1178 (depth at least
// iDepth; RebuildTopology while a full rebuild is pending; else the
// incremental path with accumulation).  A depth above the root's level
// returns -6: the reference would walk past its root.
static int accumulateTopology(Vol& v, const float* pts, int n, float radius, const float orig[3], int iDepth) {
    boundingBox(v, pts, n, orig);
    int rootPos[3];
    int depth = determineDepth(v, rootPos);
    if (depth < iDepth) depth = iDepth;
    if (depth == 0) return -4;
    if (depth > 5 || depth >= v.levels) return -4;
    if (v.rebuild_topo) return rebuildTopology(v, pts, n, radius, orig);
    if (v.root == ID_UNDEFL || depth > ElemLev(v.root)) return -6;
    v.undefined_events = 0;
    activateIncre(v, pts, n, radius, orig, depth, true);
    return 0;
}

static bool getPosInNode(Vol& v, u64 cur, const int pos[3], uint32_t& bit) {
    // src/gvdb_volume_gvdb.cpp:2858-2875
    Node* c = getNode(v, cur);
    int lev = c->mLev;
    i64 rg = v.range[lev];
    int res = v.res[lev];
    i64 p[3] = {pos[0] - (i64)c->mPos[0], pos[1] - (i64)c->mPos[1], pos[2] - (i64)c->mPos[2]};
    if (p[0] >= 0 && p[1] >= 0 && p[2] >= 0 && p[0] < rg && p[1] < rg && p[2] < rg) {
        for (int a = 0; a < 3; a++) p[a] = (p[a] * res) / rg;
        bit = uint32_t((p[2] * res + p[1]) * res + p[0]);
        return true;
    }
    bit = 0;
    return false;
}

static u64 getChildNodeH(Vol& v, u64 nodeid, uint32_t b) {
    // src/gvdb_volume_gvdb.cpp:3079-3092
    Node* c = getNode(v, nodeid);
    if (c->mChildList == ID_UNDEFL) return ID_UNDEF64;
    return PoolData64(v, c->mChildList)[b];
}

static u64 InsertChild(Vol& v, u64 nodeid, u64 childid, uint32_t i) {
    // src/gvdb_volume_gvdb.cpp:2980-3027 (dense child list branch)
    Node* cur = getNode(v, nodeid);
    Node* ch = getNode(v, childid);
    ch->mParent = nodeid;
    if (cur->mChildList == ID_UNDEFL) {
        u64 cl = PoolAlloc(v, 1, cur->mLev);
        cur = getNode(v, nodeid);
        cur->mChildList = cl;
        memset(PoolData64(v, cl), 0xFF, sizeof(u64) * voxcnt(v, cur->mLev));
    }
    PoolData64(v, cur->mChildList)[i] = childid;
    return childid;
}

static u64 AddChildNode(Vol& v, u64 nodeid, const int ppos[3], int plev, uint32_t i) {
    // src/gvdb_volume_gvdb.cpp:2707-2723, getPosFromBit gvdb_volume_gvdb.h:602-609
    u64 child = PoolAlloc(v, 0, plev - 1);
    i64 rg = v.range[plev - 1];
    int logr = v.logdim[plev];
    uint32_t mask = (1u << logr) - 1;
    int lp[3] = {int(i & mask), int((i >> logr) & mask), int((i >> (2 * logr)) & mask)};
    int p[3] = {int(lp[0] * rg + ppos[0]), int(lp[1] * rg + ppos[1]), int(lp[2] * rg + ppos[2])};
    SetupNode(v, child, plev - 1, p);
    return InsertChild(v, nodeid, child, i);
}

static u64 ActivateSpaceR(Vol& v, u64 nodeid, const int pos[3], bool& bNew, u64 stopnode, int stoplev);

static u64 Reparent(Vol& v, int lev, u64 prevroot_id, const int pos[3], bool& bNew) {
    // src/gvdb_volume_gvdb.cpp:2726-2763
    Node* pr = getNode(v, prevroot_id);
    int prpos[3] = {pr->mPos[0], pr->mPos[1], pr->mPos[2]};
    int p1[3] = {0, 0, 0}, p2[3];
    bool cover = false;
    while (!cover && lev < MAXLEV_HOST) {
        lev++;
        if (lev >= v.levels) { lev = MAXLEV_HOST; break; }
        GetCoveringNodeHost(v, lev, pos, p1);
        GetCoveringNodeHost(v, lev, prpos, p2);
        cover = (p1[0] == p2[0] && p1[1] == p2[1] && p1[2] == p2[2]);
    }
    if (lev >= MAXLEV_HOST) return ID_UNDEFL;
    u64 nr = PoolAlloc(v, 0, lev);
    if (nr == ID_UNDEFL) return ID_UNDEFL;
    SetupNode(v, nr, lev, p1);
    bool bn = false;
    ActivateSpaceR(v, nr, prpos, bn, prevroot_id, 0);
    u64 leaf = ActivateSpaceR(v, nr, pos, bNew, ID_UNDEFL, 0);
    v.root = nr;
    return getNode(v, leaf)->mParent;
}

static u64 ActivateSpaceR(Vol& v, u64 nodeid, const int pos[3], bool& bNew, u64 stopnode, int stoplev) {
    // src/gvdb_volume_gvdb.cpp:2804-2855
    if (v.root == ID_UNDEFL && nodeid == v.root) {
        int p[3];
        GetCoveringNodeHost(v, stoplev, pos, p);
        v.root = PoolAlloc(v, 0, stoplev);
        SetupNode(v, v.root, stoplev, p);
        nodeid = v.root;
    }
    uint32_t b;
    if (getPosInNode(v, nodeid, pos, b)) {
        Node* cur = getNode(v, nodeid);
        if (stopnode != ID_UNDEFL) {
            Node* sn = getNode(v, stopnode);
            if (pos[0] == sn->mPos[0] && pos[1] == sn->mPos[1] && pos[2] == sn->mPos[2] &&
                getChildNodeH(v, nodeid, b) == ID_UNDEF64 && cur->mLev == sn->mLev + 1)
                return InsertChild(v, nodeid, stopnode, b);
        }
        if (cur->mLev == stoplev) return nodeid;
        u64 child;
        if (getChildNodeH(v, nodeid, b) == ID_UNDEF64) {
            int cp[3] = {cur->mPos[0], cur->mPos[1], cur->mPos[2]};
            int clev = cur->mLev;
            child = AddChildNode(v, nodeid, cp, clev, b);
            if (clev == 1) bNew = true;
        } else {
            child = getChildNodeH(v, nodeid, b);
        }
        if (ElemLev(child) == 0) return child;
        return ActivateSpaceR(v, child, pos, bNew, stopnode, stoplev);
    } else {
        u64 parent = getNode(v, nodeid)->mParent;
        if (parent == ID_UNDEFL) {
            parent = Reparent(v, getNode(v, nodeid)->mLev, nodeid, pos, bNew);
            if (parent == ID_UNDEFL) return ID_UNDEFL;
        }
        return ActivateSpaceR(v, parent, pos, bNew, stopnode, stoplev);
    }
}

// ------------------------------------------------------- finish & atlas ----
static void computeBounds(Vol& v) {
    // src/gvdb_volume_gvdb.cpp:1792-1816
    i64 rg = v.range[0];
    Node* c = nodeAt(v, 0, 0);
    float mn[3] = {float(c->mPos[0]), float(c->mPos[1]), float(c->mPos[2])};
    float mx[3] = {mn[0], mn[1], mn[2]};
    for (u64 n = 0; n < v.pool[0][0].lastEle; n++) {
        c = nodeAt(v, 0, n);
        if (!c->mFlags) continue;
        for (int a = 0; a < 3; a++) {
            if (c->mPos[a] < mn[a]) mn[a] = float(c->mPos[a]);
            if (c->mPos[a] + rg > mx[a]) mx[a] = float(c->mPos[a] + rg);
        }
    }
    for (int a = 0; a < 3; a++) { v.objMin[a] = mn[a]; v.objMax[a] = mx[a]; }
}

static inline int brickres(const Vol& v) { return v.brickwid + 2 * v.apron; }
static inline void atlasRes(const Vol& v, int r[3]) {
    for (int a = 0; a < 3; a++) r[a] = v.atlas_cnt[a] * brickres(v);
}

static void atlasResizeTo(Vol& v, int cz) {
    // src/gvdb_allocator.cpp:622-650 (Z-axis growth, preserving the old content)
    if (cz == v.atlas_cnt[2]) return;
    v.atlas_cnt[2] = cz;
    v.atlas_max = u64(v.atlas_cnt[0]) * v.atlas_cnt[1] * v.atlas_cnt[2];
    int r[3];
    atlasRes(v, r);
    size_t sz = size_t(r[0]) * r[1] * r[2];
    v.atlasF.resize(sz, 0.f);
    if (v.fp64) v.atlasD.resize(sz, 0.0);
    if (v.has_clr) v.atlasC.resize(sz, 0u);
}

static void updateAtlas(Vol& v) {
    // src/gvdb_volume_gvdb.cpp:2630-2704
    u64 total = v.pool[0][0].lastEle, used = v.pool[0][0].usedNum;
    v.atlas_last = v.atlas_used = 0;
    if (used > v.atlas_max) {
        int cz = int(ceil(used / float(v.atlas_cnt[0] * v.atlas_cnt[1])));
        atlasResizeTo(v, cz);
    }
    int br = brickres(v);
    for (u64 n = 0; n < total; n++) {
        Node* nd = nodeAt(v, 0, n);
        if (!nd->mFlags) continue;
        if (v.atlas_last >= v.atlas_max) {
            u64 layer = u64(v.atlas_cnt[0]) * v.atlas_cnt[1];
            atlasResizeTo(v, int(ceil((v.atlas_last + layer) / float(layer))));
        }
        u64 id = v.atlas_last++;
        v.atlas_used++;
        // getAtlasPos: src/gvdb_allocator.cpp:705-715
        u64 a2 = u64(v.atlas_cnt[0]) * v.atlas_cnt[1];
        int pz = int(id / a2); id -= u64(pz) * a2;
        int py = int(id / v.atlas_cnt[0]); id -= u64(py) * v.atlas_cnt[0];
        int px = int(id);
        nd->mValue[0] = px * br + v.apron;
        nd->mValue[1] = py * br + v.apron;
        nd->mValue[2] = pz * br + v.apron;
    }
    // AllocateAtlasMap + ClearMapping + AssignMapping: src/gvdb_allocator.cpp:471-496, gvdb_volume_gvdb.cpp:2559-2588
    v.amap_max = u64(v.atlas_cnt[0]) * v.atlas_cnt[1] * v.atlas_cnt[2];
    v.amap.assign(v.amap_max, AtlasNode{{-1, -1, -1}, -1});
    for (u64 n = 0; n < total; n++) {
        Node* nd = nodeAt(v, 0, n);
        if (!nd->mFlags) continue;
        int bi[3] = {nd->mValue[0] / br, nd->mValue[1] / br, nd->mValue[2] / br};
        u64 id = (u64(bi[2]) * v.atlas_cnt[1] + bi[1]) * v.atlas_cnt[0] + bi[0];
        AtlasNode& an = v.amap[id];
        an.mPos[0] = nd->mPos[0]; an.mPos[1] = nd->mPos[1]; an.mPos[2] = nd->mPos[2];
        an.mLeafNode = int(n);
    }
    computeTopLev(v);
}

static inline size_t aidx(const Vol& v, int x, int y, int z) {
    int r[3];
    atlasRes(v, r);
    return (size_t(z) * r[1] + y) * r[0] + x;
}

// ------------------------------------------------ InsertPointsSubcell ----
// src/gvdb_volume_gvdb.cpp:4971-5092; kernels/cuda_gvdb_particles.cuh:109-118
// (SetFlagSubcell), :275-315 (CountSubcell), :218-273 (InsertSubcell),
// :677-694 (CalcSubcellPos).  Entries inside a subcell are kept in ascending
// point order (the canonical order; the reference's atomics leave it unordered).
// rule 0: InsertPointsSubcell's binning; rule 1: InsertPointsSubcell_FP16's
// (gvdb_volume_gvdb.cpp:5094-5189; cuda_gvdb_particles.cuh:154-216), whose
// insert pass tests p + trans for NOHIT and for < 0, where the float
// variant's tests p < 0.  A point is binned when its count pass and its insert
// pass both take it.
static int insertPointsSubcell(Vol& v, int scs, const float* pts, int n, float radius, const float trans[3], int* outLen,
                               int rule = 0, const uint32_t* clr = nullptr) {
    int numBrick = int(v.pool[0][0].usedNum);
    *outLen = 0;
    v.sc16 = false;
    if (numBrick == 0) return -5;
    int res0 = v.res[0];
    int spb = (res0 / scs) * (res0 / scs) * (res0 / scs);
    int numSC = spb * numBrick;
    int scdim = res0 / scs;
    int numMap = int(v.pool[0][0].lastEle);
    v.sc_mapping.assign(numMap, 0);
    {
        int acc = 0;
        for (int i = 0; i < numMap; i++) {
            v.sc_mapping[i] = acc;
            acc += (nodeAt(v, 0, i)->mParent == ID_UNDEF64) ? 0 : 1;
        }
    }
    v.num_sc = numSC;
    v.sc_cnt.assign(numSC, 0);
    auto forEachSubcell = [&](const float* wp, auto&& fn) {
        int scmin[3], scmax[3];
        for (int a = 0; a < 3; a++) {
            scmin[a] = (int(wp[a] - radius) / scs) * scs;
            scmax[a] = (int(wp[a] + 1 + radius) / scs) * scs;
        }
        int sp[3];
        for (sp[0] = scmin[0]; sp[0] <= scmax[0]; sp[0] += scs)
            for (sp[1] = scmin[1]; sp[1] <= scmax[1]; sp[1] += scs)
                for (sp[2] = scmin[2]; sp[2] <= scmax[2]; sp[2] += scs) {
                    int pid = getPosLeafParentD(v, sp);
                    if (pid < 0) continue;
                    Node* pn = nodeAt(v, 0, u64(pid));
                    int pin[3] = {(sp[0] - pn->mPos[0]) / scs, (sp[1] - pn->mPos[1]) / scs, (sp[2] - pn->mPos[2]) / scs};
                    // a top node that is a leaf is found for corners past it too,
                    // whose offsets index past every subcell table
                    if (pin[0] >= scdim || pin[1] >= scdim || pin[2] >= scdim) { v.undefined_events++; continue; }
                    int lo = (pin[2] * scdim + pin[1]) * scdim + pin[0];
                    // a corner less than a cell below the brick's (truncated
                    // into its first cell) gives a negative offset: before the
                    // first linked brick that indexes outside sc_cnt
                    const i64 idx = i64(v.sc_mapping[pid]) * spb + lo;
                    if (idx < 0) { v.undefined_events++; continue; }
                    fn(int(idx));
                }
    };
    // The reference's count pass tests the position after adding trans, its
    // insert pass before; when the two disagree it writes past the counted
    // entries (undefined).  Such points are recorded and skipped in both.
    auto binned = [&](const float* p, float w[3]) {
        if (p[2] == NOHIT) return false;
        for (int a = 0; a < 3; a++) w[a] = p[a] + trans[a];
        bool wneg = w[0] < 0 || w[1] < 0 || w[2] < 0, pneg = p[0] < 0 || p[1] < 0 || p[2] < 0;
        if (rule == 1) return !wneg && w[2] != NOHIT;
        return !wneg && !pneg;
    };
    // Both passes run over the same static chunks of points, one per thread:
    // per-thread counts per subcell, and per-thread starting slots from them,
    // so each subcell's entries stay in ascending point order.
    const int nt = std::max(1, std::min(omp_get_max_threads(), n / 4096 + 1));
    std::vector<std::vector<int>> tc(nt, std::vector<int>(numSC, 0));
    i64 undef = 0;
#pragma omp parallel for num_threads(nt) schedule(static, 1) reduction(+ : undef)
    for (int t = 0; t < nt; t++) {
        const int i0 = int(i64(n) * t / nt), i1 = int(i64(n) * (t + 1) / nt);
        std::vector<int>& c = tc[t];
        for (int i = i0; i < i1; i++) {
            const float* p = pts + 3 * i;
            float w[3];
            if (p[2] == NOHIT) continue;
            bool wneg = p[0] + trans[0] < 0 || p[1] + trans[1] < 0 || p[2] + trans[2] < 0;
            bool pneg = p[0] < 0 || p[1] < 0 || p[2] < 0;
            if (rule == 0 && wneg != pneg) undef++;
            if (rule == 1 && !wneg && p[2] + trans[2] == NOHIT) undef++;
            if (!binned(p, w)) continue;
            forEachSubcell(w, [&](int idx) { c[idx]++; });
        }
    }
    v.undefined_events += int(undef);
    v.sc_off.assign(numSC, 0);
    {
        i64 acc = 0;
        for (int i = 0; i < numSC; i++) {
            v.sc_off[i] = int(acc);
            int tot = 0;
            for (int t = 0; t < nt; t++) {
                const int c = tc[t][i];
                tc[t][i] = int(acc) + tot;      // thread t's first slot in subcell i
                tot += c;
            }
            v.sc_cnt[i] = tot;
            acc += tot;
        }
    }
    int len = v.sc_off[numSC - 1] + v.sc_cnt[numSC - 1];
    *outLen = len;
    v.sc_len = len;
    // insert pass
    v.sc_pnt.assign(size_t(len) * 3, 0.f);
    // InsertSubcell's sc_pnt_clr (kernels/cuda_gvdb_particles.cuh:244,269)
    if (clr) v.sc_clr.assign(size_t(len), 0u);
    else v.sc_clr.clear();
    if (len > 0) {
#pragma omp parallel for num_threads(nt) schedule(static, 1)
        for (int t = 0; t < nt; t++) {
            const int i0 = int(i64(n) * t / nt), i1 = int(i64(n) * (t + 1) / nt);
            std::vector<int>& slot = tc[t];
            for (int i = i0; i < i1; i++) {
                const float* p = pts + 3 * i;
                float w[3];
                if (!binned(p, w)) continue;
                forEachSubcell(w, [&](int idx) {
                    size_t e = size_t(slot[idx]++);
                    v.sc_pnt[3 * e] = w[0]; v.sc_pnt[3 * e + 1] = w[1]; v.sc_pnt[3 * e + 2] = w[2];
                    if (clr) v.sc_clr[e] = clr[i];
                });
            }
        }
    }
    // CalcSubcellPos.  The reference returns before it on an empty list
    // (gvdb_volume_gvdb.cpp:5058), which leaves GatherLevelSet reading the
    // corners of an earlier insert or unallocated tables; here an empty list
    // still gets its corners, so the level set is 3 on every linked brick.
    v.sc_pos.assign(size_t(numSC) * 3, 0);
    v.sc_nid.assign(numSC, 0);
    for (int i = 0; i < numMap; i++) {
        Node* pn = nodeAt(v, 0, i);
        if (pn->mParent == ID_UNDEF64) continue;
        for (int s = 0; s < spb; s++) {
            size_t idx = size_t(v.sc_mapping[i]) * spb + s;
            v.sc_pos[3 * idx] = pn->mPos[0] + s % scdim * scs;
            v.sc_pos[3 * idx + 1] = pn->mPos[1] + (s / scdim) % scdim * scs;
            v.sc_pos[3 * idx + 2] = pn->mPos[2] + (s / scdim / scdim) % scdim * scs;
            v.sc_nid[idx] = i;
        }
    }
    return 0;
}

// ------------------------------------------------------ GatherDensity ----
// src/gvdb_volume_gvdb.cpp:5318-5344; kernels/cuda_gvdb_particles.cuh:696-733
// One subcell at a time, each voxel takes max(val, radius - |p - v|) over the
// subcell's points, v being the integer voxel coordinate.
// levelset = false: gvdbGatherDensity (kernels/cuda_gvdb_particles.cuh:696-733),
//   val = max(val, r - c) from 0 (or the atlas value when accumulating);
// levelset = true: gvdbGatherLevelSet (kernels/cuda_gvdb_particles.cuh:736-774),
//   dist = min(dist, c - r) from 3 (or the atlas value when accumulating).
// gvdbGatherLevelSet_fp16 (cuda_gvdb_particles.cuh:776-811) reads an entry
// coordinate as u / 65535.0f * range + min; the product and the sum are
// fused, as a contracting compiler emits them.
static float dequant16(uint16_t u, float mn, float rg) { return std::fmaf(float(u) / 65535.0f, rg, mn); }

// fp16: the entries are InsertPointsSubcell_FP16's u16 ones (level set only).
template <typename T>
static void gatherT(Vol& v, int scs, float radius, bool accum, std::vector<T>& atlas, bool levelset = false,
                    bool fp16 = false) {
    int spb = (v.res[0] / scs) * (v.res[0] / scs) * (v.res[0] / scs);
    int numSC = spb * int(v.pool[0][0].usedNum);
    const T R = T(radius);
#pragma omp parallel for schedule(dynamic, 64)
    for (int s = 0; s < numSC; s++) {
        Node* nd = nodeAt(v, 0, u64(v.sc_nid[s]));
        for (int tz = 0; tz < scs; tz++)
            for (int ty = 0; ty < scs; ty++)
                for (int tx = 0; tx < scs; tx++) {
                    int w[3] = {v.sc_pos[3 * s] + tx, v.sc_pos[3 * s + 1] + ty, v.sc_pos[3 * s + 2] + tz};
                    int vox[3] = {nd->mValue[0] + (w[0] - nd->mPos[0]), nd->mValue[1] + (w[1] - nd->mPos[1]),
                                  nd->mValue[2] + (w[2] - nd->mPos[2])};
                    size_t ai = aidx(v, vox[0], vox[1], vox[2]);
                    T val = accum ? atlas[ai] : (levelset ? T(3) : T(0));
                    size_t off = size_t(v.sc_off[s]);
                    for (int j = 0; j < v.sc_cnt[s]; j++) {
                        float pq[3];
                        const float* p = &v.sc_pnt[3 * (off + j)];
                        if (fp16) {
                            for (int a = 0; a < 3; a++) pq[a] = dequant16(v.sc_pnt16[3 * (off + j) + a], v.qmin[a], v.qrange[a]);
                            p = pq;
                        }
                        T jx = T(p[0]) - T(w[0]), jy = T(p[1]) - T(w[1]), jz = T(p[2]) - T(w[2]);
                        T c = std::sqrt(jx * jx + jy * jy + jz * jz);
                        val = levelset ? std::min(val, c - R) : std::max(val, R - c);
                    }
                    atlas[ai] = val;
                }
    }
}

// -------------------------------------------------- colour of the gather ----
// gvdbGatherDensity / gvdbGatherLevelSet with a colour table
// (kernels/cuda_gvdb_particles.cuh:717-731, :757-771): per voxel of a subcell,
// clr = (0, 0, 0, 1), or CHAR2CLR(the voxel's colour) with bAccumulate; add
// CHAR2CLR(entry colour) = byte / 255.0f per component for every entry of the
// subcell's list; divide by the entry count (+1 with bAccumulate); store
// CLR2CHAR = uchar(x * 255.0f) per component.  The float-to-uchar conversion
// truncates and saturates (NaN: 0), so an empty list stores (0, 0, 0, 255).
// exact = false: that float arithmetic in list order (which the reference's
// atomics leave unordered); exact = true: the same mean taken exactly,
// (255 [alpha only] + sum of bytes) / count, or (old + sum) / (count + 1), in
// integer division.
static uint8_t toU8(float x) {
    if (!(x > 0.f)) return 0;                  // NaN and non-positive
    if (x >= 255.f) return 255;
    return uint8_t(x);
}
static void gatherColor(Vol& v, int scs, bool accum, bool exact) {
    int spb = (v.res[0] / scs) * (v.res[0] / scs) * (v.res[0] / scs);
    int numSC = spb * int(v.pool[0][0].usedNum);
#pragma omp parallel for schedule(dynamic, 64)
    for (int s = 0; s < numSC; s++) {
        Node* nd = nodeAt(v, 0, u64(v.sc_nid[s]));
        const int cnt = v.sc_cnt[s];
        const size_t off = size_t(v.sc_off[s]);
        uint64_t sum[4] = {0, 0, 0, 0};
        for (int j = 0; j < cnt; j++) {
            const uint32_t c = v.sc_clr[off + j];
            for (int k = 0; k < 4; k++) {
                sum[k] += (c >> (8 * k)) & 0xFFu;
            }
        }
        for (int tz = 0; tz < scs; tz++)
            for (int ty = 0; ty < scs; ty++)
                for (int tx = 0; tx < scs; tx++) {
                    int w[3] = {v.sc_pos[3 * s] + tx, v.sc_pos[3 * s + 1] + ty, v.sc_pos[3 * s + 2] + tz};
                    int vox[3] = {nd->mValue[0] + (w[0] - nd->mPos[0]), nd->mValue[1] + (w[1] - nd->mPos[1]),
                                  nd->mValue[2] + (w[2] - nd->mPos[2])};
                    size_t ai = aidx(v, vox[0], vox[1], vox[2]);
                    const uint32_t old = v.atlasC[ai];
                    uint32_t out = 0;
                    for (int k = 0; k < 4; k++) {
                        const uint32_t ob = (old >> (8 * k)) & 0xFFu;
                        uint32_t b;
                        if (exact) {
                            if (accum) b = uint32_t((uint64_t(ob) + sum[k]) / uint64_t(cnt + 1));
                            else if (cnt == 0) b = k == 3 ? 255u : 0u;
                            else b = uint32_t(std::min<uint64_t>((sum[k] + (k == 3 ? 255u : 0u)) / uint64_t(cnt), 255u));
                        } else {
                            float f = accum ? float(ob) / 255.0f : (k == 3 ? 1.0f : 0.0f);
                            for (int j = 0; j < cnt; j++) f += float((v.sc_clr[off + j] >> (8 * k)) & 0xFFu) / 255.0f;
                            f /= float(cnt + (accum ? 1 : 0));
                            b = toU8(f * 255.0f);
                        }
                        out |= b << (8 * k);
                    }
                    v.atlasC[ai] = out;
                }
    }
}

// -------------------------------------------------------- UpdateApron ----
// src/gvdb_volume_gvdb.cpp:4418-4457; kernels/cuda_gvdb_operators.cuh:72-131
// with getAtlasToWorld (kernels/cuda_gvdb_nodes.cuh:145-153) and getNodeAtPoint.
// Unflagged leaves (left by the incremental build) are skipped: the reference
// runs them too, through an atlas position that is either stale, and then
// rewrites the value its current owner writes, or (-1, -1, -1), which it turns
// into an out-of-range atlas-map read.
// T: the voxel type; C: the coordinate arithmetic (UpdateApron<uchar4>,
// kernels/cuda_gvdb_operators.cuh:143-146, walks in float like the float one).
template <typename T, typename C = T>
static void apronT(Vol& v, T boundval, std::vector<T>& atlas) {
    int bricks = int(v.pool[0][0].lastEle);
    int pbr = brickres(v);
#pragma omp parallel for schedule(dynamic, 16)
    for (int b = 0; b < bricks; b++) {
        Node* node = nodeAt(v, 0, b);
        if (!node->mFlags) continue;
        for (int side = 0; side < 6; side++)
            for (int sy = 0; sy < pbr; sy++)
                for (int sx = 0; sx < pbr; sx++) {
                    int bv[3];
                    switch (side) {
                        case 0: bv[0] = 0; bv[1] = sx; bv[2] = sy; break;
                        case 1: bv[0] = sx; bv[1] = 0; bv[2] = sy; break;
                        case 2: bv[0] = sx; bv[1] = sy; bv[2] = 0; break;
                        case 3: bv[0] = pbr - 1; bv[1] = sx; bv[2] = sy; break;
                        case 4: bv[0] = sx; bv[1] = pbr - 1; bv[2] = sy; break;
                        default: bv[0] = sx; bv[1] = sy; bv[2] = pbr - 1; break;
                    }
                    int av[3] = {bv[0] + node->mValue[0] - 1, bv[1] + node->mValue[1] - 1, bv[2] + node->mValue[2] - 1};
                    int bi[3] = {av[0] / pbr, av[1] / pbr, av[2] / pbr};
                    u64 bid = (u64(bi[2]) * v.atlas_cnt[1] + bi[1]) * v.atlas_cnt[0] + bi[0];
                    const AtlasNode& an = v.amap[bid];
                    if (an.mLeafNode == -1) continue;
                    C wp[3];
                    for (int a = 0; a < 3; a++) wp[a] = C(an.mPos[a]) + C(av[a] % pbr - v.apron) + C(0.5);
                    i64 id = getNodeAtPointT<C>(v, wp);
                    T value;
                    if (id < 0) value = boundval;
                    else {
                        Node* nn = nodeAt(v, 0, u64(id));
                        int o[3];
                        for (int a = 0; a < 3; a++) o[a] = int(unsigned(C(nn->mValue[a]) + (wp[a] - C(nn->mPos[a]))));
                        value = atlas[aidx(v, o[0], o[1], o[2])];
                    }
                    atlas[aidx(v, av[0], av[1], av[2])] = value;
                }
    }
}

// ------------------------------------------------------------- camera ----
// Camera3D::setOrbit/updateMatricies/updateFrustum/inverseRayProj:
// src/gvdb_camera.cpp:90-105, :175-228, :292-346, :382-393; matrix helpers
// src/gvdb_vec.cpp:375-392 (Basis), :427-466 (InvertTRS), :487-495
// (PreTranslate), :513-520 (InverseProj), :522-530 (InverseView).
struct Mat4 { float d[16]; };
static Mat4 mmul(const Mat4& a, const Mat4& b) {
    // Matrix4F::operator*= (column-major): result = a * b
    Mat4 r;
    for (int c = 0; c < 4; c++)
        for (int rr = 0; rr < 4; rr++) {
            float s = 0;
            for (int k = 0; k < 4; k++) s += a.d[k * 4 + rr] * b.d[c * 4 + k];
            r.d[c * 4 + rr] = s;
        }
    return r;
}
static void invertGeneral(Mat4& m) {
    double a[16], inv[16];
    for (int i = 0; i < 16; i++) a[i] = m.d[i];
    // cofactor expansion (any correct 4x4 inverse); computed in double as the reference does
    inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    if (det == 0) return;
    det = 1.0 / det;
    for (int i = 0; i < 16; i++) m.d[i] = float(inv[i] * det);
}

// Scene block shared with the product (layout documented in include/gvdb_amd.h)
struct ScnBlock {
    int32_t width, height;
    float campos[3], cams[3], camu[3], camv[3];
    float light_pos[3];
    float backclr[4];
    float extinct[3], steps[3], cutoff[3], thresh[3], shadow[3];
    float invxform[16], invxrot[16];
};

struct CamParams {
    float angs[3], to[3], dist, dolly, fov, znear, zfar;
};

static void cameraRays(const CamParams& cp, int w, int h, float campos[3], float tl[3], float tr[3], float bl[3]) {
    const float D2R = 3.141592f / 180.0f;   // DEGtoRAD: src/gvdb_types.h:59
    float from[3];
    double dx = cos(double(cp.angs[1] * D2R)) * sin(double(cp.angs[0] * D2R));
    double dy = sin(double(cp.angs[1] * D2R));
    double dz = cos(double(cp.angs[1] * D2R)) * cos(double(cp.angs[0] * D2R));
    from[0] = cp.to[0] + float(dx) * cp.dist;
    from[1] = cp.to[1] + float(dy) * cp.dist;
    from[2] = cp.to[2] + float(dz) * cp.dist;
    float aspect = float(w) / float(h);   // Scene::SetRes -> SetAspect (src/gvdb_scene.cpp:399-404)
    // gluLookAt basis: Vector3D::Cross in double then cast to float, Normalize in double (src/gvdb_vec.h:120-130, :256-264)
    auto normf = [](float* v) {
        double n = double(v[0]) * v[0] + double(v[1]) * v[1] + double(v[2]) * v[2];
        if (n != 0.0) { double r = 1.0 / sqrt(n); for (int a = 0; a < 3; a++) v[a] = float(v[a] * r); }
    };
    auto crossf = [](const float* a, const float* b, float* o) {
        double ax = a[0], ay = a[1], az = a[2];
        o[0] = float(ay * double(b[2]) - az * double(b[1]));
        o[1] = float(-ax * double(b[2]) + az * double(b[0]));
        o[2] = float(ax * double(b[1]) - ay * double(b[0]));
    };
    float fdir[3] = {cp.to[0] - from[0], cp.to[1] - from[1], cp.to[2] - from[2]};
    normf(fdir);
    const float upd[3] = {0, 1, 0};
    float fside[3], fup[3];
    crossf(fdir, upd, fside);
    normf(fside);
    crossf(fside, fdir, fup);
    normf(fup);
    float fz[3] = {-fdir[0], -fdir[1], -fdir[2]};
    Mat4 rot;   // Basis(side, up, dir): rows are the basis vectors
    float b[16] = {fside[0], fup[0], fz[0], 0, fside[1], fup[1], fz[1], 0, fside[2], fup[2], fz[2], 0, 0, 0, 0, 1};
    memcpy(rot.d, b, sizeof(b));
    Mat4 view = rot;   // translation removed below anyway
    float sx = float(tan(cp.fov * D2R / 2.0f)) * cp.znear;
    float sy = sx / aspect;
    Mat4 proj;
    memset(proj.d, 0, sizeof(proj.d));
    proj.d[0] = 2.0f * cp.znear / sx;
    proj.d[5] = 2.0f * cp.znear / sy;
    proj.d[10] = -(cp.zfar + cp.znear) / (cp.zfar - cp.znear);
    proj.d[14] = -(2.0f * cp.zfar * cp.znear) / (cp.zfar - cp.znear);
    proj.d[11] = -1.0f;
    view.d[12] = view.d[13] = view.d[14] = 0;
    Mat4 ivp = mmul(proj, view);
    invertGeneral(ivp);
    auto irp = [&](float x, float y, float z, float out[3]) {
        float p[4] = {x, y, z, 1.0f};
        float wv[4];
        for (int r = 0; r < 4; r++) wv[r] = ivp.d[r] * p[0] + ivp.d[4 + r] * p[1] + ivp.d[8 + r] * p[2] + ivp.d[12 + r];
        out[0] = wv[0] / wv[3]; out[1] = wv[1] / wv[3]; out[2] = wv[2] / wv[3];
    };
    irp(-1.0f, 1.0f, cp.znear, tl);
    irp(1.0f, 1.0f, cp.znear, tr);
    irp(-1.0f, -1.0f, cp.znear, bl);
    campos[0] = from[0]; campos[1] = from[1]; campos[2] = from[2];
}

// ------------------------------------------------------------- render ----
template <typename T> struct V3 { T x, y, z; };
template <typename T> static inline V3<T> mk(T a, T b, T c) { return V3<T>{a, b, c}; }
template <typename T> static inline V3<T> operator+(V3<T> a, V3<T> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T> static inline V3<T> operator-(V3<T> a, V3<T> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T> static inline V3<T> operator*(V3<T> a, T s) { return mk(a.x * s, a.y * s, a.z * s); }
template <typename T> static inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> static inline V3<T> normalize(V3<T> a) { T il = T(1) / std::sqrt(dot(a, a)); return a * il; }

template <typename T> struct RenderCtx {
    Vol* v;
    const std::vector<T>* atlas;
    int ares[3];
    const float* tf;   // 16384 x float4
    const std::vector<uint32_t>* clr = nullptr;   // SetColorChannel's uchar4 channel, or none
    V3<T> light;
    T extinct, albedo, dstep, minval, alphacut, thresh, vmin, vmax, shadowamt, shadowbias, eps;
    T back[4];
};

// Unnormalised linear texture fetch with clamp addressing (exact weights),
// blended as a + f (b - a) along x, then y, then z.  The reference leaves the
// weights to the texture unit (8-bit fixed-point fractions), so the blend
// form is not the reference's; this one is fixed so that an FP32 march can be
// compared operation for operation.
template <typename T> static T tex3(const RenderCtx<T>& c, T x, T y, T z) {
    T u[3] = {x - T(0.5), y - T(0.5), z - T(0.5)};
    int i0[3], i1[3];
    T f[3];
    for (int a = 0; a < 3; a++) {
        T fl = std::floor(u[a]);
        f[a] = u[a] - fl;
        int i = int(fl);
        i0[a] = std::min(std::max(i, 0), c.ares[a] - 1);
        i1[a] = std::min(std::max(i + 1, 0), c.ares[a] - 1);
    }
    auto at = [&](int X, int Y, int Z) { return (*c.atlas)[(size_t(Z) * c.ares[1] + Y) * c.ares[0] + X]; };
    auto lerp = [](T a, T b, T w) { return a + w * (b - a); };
    T c00 = lerp(at(i0[0], i0[1], i0[2]), at(i1[0], i0[1], i0[2]), f[0]);
    T c10 = lerp(at(i0[0], i1[1], i0[2]), at(i1[0], i1[1], i0[2]), f[0]);
    T c01 = lerp(at(i0[0], i0[1], i1[2]), at(i1[0], i0[1], i1[2]), f[0]);
    T c11 = lerp(at(i0[0], i1[1], i1[2]), at(i1[0], i1[1], i1[2]), f[0]);
    T c0 = lerp(c00, c10, f[1]);
    T c1 = lerp(c01, c11, f[1]);
    return lerp(c0, c1, f[2]);
}

template <typename T> struct HDDA {
    // kernels/cuda_gvdb_dda.cuh:38-91
    V3<T> pos, dir, tDel, t, tSide;
    int pStep[3], p[3], mask[3];
    void SetFromRay(V3<T> sp, V3<T> sd, V3<T> st) {
        pos = sp; dir = sd; t = st;
        pStep[0] = dir.x > 0 ? 1 : -1; pStep[1] = dir.y > 0 ? 1 : -1; pStep[2] = dir.z > 0 ? 1 : -1;
    }
    void Prepare(V3<T> vmin, T vdel) {
        tDel = mk(std::fabs(vdel / dir.x), std::fabs(vdel / dir.y), std::fabs(vdel / dir.z));
        V3<T> pf = (pos + dir * t.x - vmin) * (T(1) / vdel);
        pf = mk((pos.x + t.x * dir.x - vmin.x) / vdel, (pos.y + t.x * dir.y - vmin.y) / vdel, (pos.z + t.x * dir.z - vmin.z) / vdel);
        T fl[3] = {std::floor(pf.x), std::floor(pf.y), std::floor(pf.z)};
        T pv[3] = {pf.x, pf.y, pf.z};
        T td[3] = {tDel.x, tDel.y, tDel.z};
        T ts[3];
        for (int a = 0; a < 3; a++) {
            ts[a] = ((fl[a] - pv[a] + T(0.5)) * T(pStep[a]) + T(0.5)) * td[a] + t.x;
            p[a] = int(fl[a]);
        }
        tSide = mk(ts[0], ts[1], ts[2]);
    }
    void Next() {
        mask[0] = int((tSide.x < tSide.y) & (tSide.x <= tSide.z));
        mask[1] = int((tSide.y < tSide.z) & (tSide.y <= tSide.x));
        mask[2] = int((tSide.z < tSide.x) & (tSide.z <= tSide.y));
        t.y = mask[0] ? tSide.x : (mask[1] ? tSide.y : tSide.z);
    }
    void Step() {
        t.x = t.y;
        tSide = tSide + mk(T(mask[0]) * tDel.x, T(mask[1]) * tDel.y, T(mask[2]) * tDel.z);
        for (int a = 0; a < 3; a++) p[a] += mask[a] * pStep[a];
    }
};

// getColorF (kernels/cuda_gvdb_raycast.cuh:206-209): the colour channel's
// voxel at the integer part of atlas position p, bytes / 255; (1, 1, 1, 1)
// without a colour channel
template <typename T> static void colorAt(const RenderCtx<T>& c, V3<T> p, T out[4]) {
    if (!c.clr) {
        for (int k = 0; k < 4; k++) out[k] = T(1);
        return;
    }
    auto cl = [&](int i, int a) { return std::min(std::max(i, 0), c.ares[a] - 1); };
    const uint32_t u = (*c.clr)[(size_t(cl(int(p.z), 2)) * c.ares[1] + cl(int(p.y), 1)) * c.ares[0] + cl(int(p.x), 0)];
    for (int k = 0; k < 4; k++) out[k] = T(float((u >> (8 * k)) & 0xFFu) / 255.0f);
}

template <typename T> static void tfLookup(const RenderCtx<T>& c, T v, T out[4]) {
    // transfer(): kernels/cuda_gvdb_dda.cuh:20-23
    T x = (v - c.thresh) / (c.vmax - c.vmin);
    x = std::max(T(0), std::min(T(1), x));
    int i = int(x * T(16300));
    for (int k = 0; k < 4; k++) out[k] = T(c.tf[4 * i + k]);
}

template <typename T>
static void rayDeepBrick(const RenderCtx<T>& c, int nodeid, V3<T> t, V3<T> pos, V3<T> dir, V3<T>& hit, T clr[4]) {
    // kernels/cuda_gvdb_raycast.cuh:485-533 (no depth buffer, no colour channel)
    Vol& v = *c.v;
    Node* nd = nodeAt(v, 0, u64(nodeid));
    V3<T> vmin = mk(T(nd->mPos[0]), T(nd->mPos[1]), T(nd->mPos[2]));
    t.x = c.dstep * std::ceil(t.x / c.dstep);
    V3<T> o = mk(T(nd->mValue[0]), T(nd->mValue[1]), T(nd->mValue[2]));
    V3<T> wp = pos + dir * t.x;
    V3<T> p = wp - vmin;
    V3<T> wpt = dir * c.dstep;
    T dt = std::sqrt(dot(wpt, wpt));
    if (hit.x == 0) hit.x = t.x;
    T r0 = T(v.res[0]);
    for (int it = 0; clr[3] > c.alphacut && it < MAX_ITER && p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < r0 && p.y < r0 && p.z < r0; it++) {
        T raw = tex3(c, p.x + o.x, p.y + o.y, p.z + o.z);
        if (raw >= c.minval) {
            T val[4];
            tfLookup(c, raw, val);
            val[3] = std::exp(c.extinct * val[3] * c.dstep);
            T h[4];
            colorAt(c, mk(p.x + o.x, p.y + o.y, p.z + o.z), h);
            clr[0] += val[0] * clr[3] * (1 - val[3]) * c.albedo * h[0];
            clr[1] += val[1] * clr[3] * (1 - val[3]) * c.albedo * h[1];
            clr[2] += val[2] * clr[3] * (1 - val[3]) * c.albedo * h[2];
            clr[3] *= val[3];
        }
        p = p + wpt;
        wp = wp + wpt;
        t.x += dt;
    }
    hit.y = t.x;
    clr[0] = std::min(clr[0], T(1)); clr[1] = std::min(clr[1], T(1)); clr[2] = std::min(clr[2], T(1));
    clr[3] = std::max(clr[3], T(0));
}

template <typename T> static V3<T> gradient(const RenderCtx<T>& c, V3<T> p) {
    // getGradient: kernels/cuda_gvdb_raycast.cuh:132-141
    V3<T> g;
    g.x = tex3(c, p.x - T(.5), p.y, p.z) - tex3(c, p.x + T(.5), p.y, p.z);
    g.y = tex3(c, p.x, p.y - T(.5), p.z) - tex3(c, p.x, p.y + T(.5), p.z);
    g.z = tex3(c, p.x, p.y, p.z - T(.5)) - tex3(c, p.x, p.y, p.z + T(.5));
    return normalize(g);
}

template <typename T>
static void raySurfTriBrick(const RenderCtx<T>& c, int nodeid, V3<T> t, V3<T> pos, V3<T> dir, V3<T>& hit, V3<T>& norm,
                            T hclr[4]) {
    // kernels/cuda_gvdb_raycast.cuh:281-300
    Vol& v = *c.v;
    Node* nd = nodeAt(v, 0, u64(nodeid));
    V3<T> vmin = mk(T(nd->mPos[0]), T(nd->mPos[1]), T(nd->mPos[2]));
    V3<T> o = mk(T(nd->mValue[0]), T(nd->mValue[1]), T(nd->mValue[2]));
    t.x = c.dstep * std::ceil(t.x / c.dstep);
    V3<T> p = pos + dir * t.x - vmin;
    T r0 = T(v.res[0]);
    for (int it = 0; it < MAX_ITER && p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < r0 && p.y < r0 && p.z < r0; it++) {
        if (tex3(c, p.x + o.x, p.y + o.y, p.z + o.z) >= c.thresh) {
            hit = p + vmin;
            norm = gradient(c, p + o);
            if (c.clr) colorAt(c, p + o, hclr);
            return;
        }
        p = p + dir * c.dstep;
        t.x += c.dstep;
    }
}

template <typename T>
static bool childOn(Vol& v, Node* node, const int p[3], int lev, int& child) {
    // the reference reads the slot even for p == res; such reads fall outside
    // the child list and are counted as undefined, the cell treated as empty
    int res = v.res[lev];
    if (p[0] < 0 || p[1] < 0 || p[2] < 0 || p[0] >= res || p[1] >= res || p[2] >= res) { v.undefined_events++; return false; }
    int b = (((p[2] << v.logdim[lev]) + p[1]) << v.logdim[lev]) + p[0];
    child = getChildD(v, node, b);
    return child >= 0;
}

// rayCast: kernels/cuda_gvdb_raycast.cuh:543-611 with rayBoxIntersect
// (kernels/cuda_gvdb_geom.cuh:85-98).  mode 0 = deep (EA), 1 = trilinear surface.
// Returns false when the walk stopped at MAX_ITER (the reference's cap, where
// the image depends on the exact step sequence).
template <typename T>
static bool rayCast(const RenderCtx<T>& c, V3<T> pos, V3<T> dir, V3<T>& hit, V3<T>& norm, T clr[4], int mode) {
    Vol& v = *c.v;
    int nodeid[8];
    T tMax[8];
    int lev = v.top_lev;
    nodeid[lev] = 0;
    T ht[8];
    ht[0] = (T(v.objMin[0]) - pos.x) / dir.x; ht[1] = (T(v.objMax[0]) - pos.x) / dir.x;
    ht[2] = (T(v.objMin[1]) - pos.y) / dir.y; ht[3] = (T(v.objMax[1]) - pos.y) / dir.y;
    ht[4] = (T(v.objMin[2]) - pos.z) / dir.z; ht[5] = (T(v.objMax[2]) - pos.z) / dir.z;
    ht[6] = std::fmax(std::fmax(std::fmin(ht[0], ht[1]), std::fmin(ht[2], ht[3])), std::fmin(ht[4], ht[5]));
    ht[7] = std::fmin(std::fmin(std::fmax(ht[0], ht[1]), std::fmax(ht[2], ht[3])), std::fmax(ht[4], ht[5]));
    ht[6] = (ht[6] < 0) ? T(0) : ht[6];
    V3<T> tStart = mk(ht[6], ht[7], (ht[7] < ht[6] || ht[7] < 0) ? T(NOHIT) : T(0));
    Node* node = nodeAt(v, lev, 0);
    if (tStart.z == T(NOHIT)) return true;
    tStart.x += c.eps;
    tMax[lev] = tStart.y - c.eps;
    HDDA<T> dda;
    dda.SetFromRay(pos, dir, tStart);
    auto vminOf = [&](Node* n) { return mk(T(n->mPos[0]), T(n->mPos[1]), T(n->mPos[2])); };
    auto vdelOf = [&](int l) { return T(v.range[l] / v.res[l]); };
    dda.Prepare(vminOf(node), vdelOf(lev));
    int it = 0;
    for (; it < MAX_ITER && lev > 0 && lev <= v.top_lev && dda.p[0] >= 0 && dda.p[1] >= 0 && dda.p[2] >= 0 &&
           dda.p[0] <= v.res[lev] && dda.p[1] <= v.res[lev] && dda.p[2] <= v.res[lev]; it++) {
        dda.Next();
        int child;
        if (childOn<T>(v, node, dda.p, lev, child)) {
            if (lev == 1) {
                dda.t.x += c.eps;
                if (mode == 0) rayDeepBrick(c, child, dda.t, pos, dir, hit, clr);
                else raySurfTriBrick(c, child, dda.t, pos, dir, hit, norm, clr);
                if (clr[3] <= 0) { clr[3] = 0; return true; }
                if (hit.z != T(NOHIT)) return true;
                dda.Step();
            } else {
                lev--;
                nodeid[lev] = child;
                node = nodeAt(v, lev, u64(child));
                dda.t.x += c.eps;
                tMax[lev] = dda.t.y - c.eps;
                dda.Prepare(vminOf(node), vdelOf(lev));
            }
        } else {
            dda.Step();
        }
        while (dda.t.x > tMax[lev] && lev <= v.top_lev) {
            lev++;
            if (lev <= v.top_lev) {
                node = nodeAt(v, lev, u64(nodeid[lev]));
                dda.Prepare(vminOf(node), vdelOf(lev));
            }
        }
    }
    return it < MAX_ITER;
}

// gvdbRayDeep / gvdbRaySurfaceTrilinear + performPhongShading:
// kernels/cuda_gvdb_module.cu:38-78, :103-120.  Output is float RGBA before
// the reference's uchar conversion.
template <typename T>
static void renderT(Vol& v, const std::vector<T>& atlas, const ScnBlock& s, const float* tf, int shading, float* out,
                    uint8_t* capped) {
    RenderCtx<T> c;
    c.v = &v;
    c.atlas = &atlas;
    atlasRes(v, c.ares);
    c.tf = tf;
    if (v.render_clr && v.has_clr) c.clr = &v.atlasC;
    c.light = mk(T(s.light_pos[0]), T(s.light_pos[1]), T(s.light_pos[2]));
    c.extinct = T(s.extinct[0]); c.albedo = T(s.extinct[1]);
    c.dstep = T(s.steps[0]);
    c.minval = T(s.cutoff[0]); c.alphacut = T(s.cutoff[1]);
    c.thresh = T(s.thresh[0]); c.vmin = T(s.thresh[1]); c.vmax = T(s.thresh[2]);
    c.shadowamt = T(s.shadow[0]); c.shadowbias = T(s.shadow[1]);
    c.eps = T(v.epsilon);
    for (int k = 0; k < 4; k++) c.back[k] = T(s.backclr[k]);
    int W = s.width, H = s.height;
    auto mm = [](const float* m, V3<T> a) {
        return mk(a.x * T(m[0]) + a.y * T(m[4]) + a.z * T(m[8]) + T(m[12]), a.x * T(m[1]) + a.y * T(m[5]) + a.z * T(m[9]) + T(m[13]),
                  a.x * T(m[2]) + a.y * T(m[6]) + a.z * T(m[10]) + T(m[14]));
    };
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            V3<T> rpos = mm(s.invxform, mk(T(s.campos[0]), T(s.campos[1]), T(s.campos[2])));
            T fx = T(float(x + 0.5f) / float(W)), fy = T(float(y + 0.5f) / float(H));
            V3<T> vv = mk(fx * T(s.camu[0]) + fy * T(s.camv[0]) + T(s.cams[0]), fx * T(s.camu[1]) + fy * T(s.camv[1]) + T(s.cams[1]),
                          fx * T(s.camu[2]) + fy * T(s.camv[2]) + T(s.cams[2]));
            V3<T> rdir = normalize(mm(s.invxrot, vv));
            T res[4];
            bool whole = true;
            if (shading == 7) {
                T clr[4] = {0, 0, 0, 1};
                V3<T> hit = mk(T(0), T(0), T(NOHIT)), norm = mk(T(0), T(0), T(0));
                whole = rayCast(c, rpos, rdir, hit, norm, clr, 0);
                T a = 1 - clr[3];
                for (int k = 0; k < 3; k++) res[k] = c.back[k] + a * (clr[k] - c.back[k]);
                res[3] = a;
            } else {
                T clr[4] = {1, 1, 1, 1};
                V3<T> hit = mk(T(NOHIT), T(NOHIT), T(NOHIT)), norm = mk(T(0), T(0), T(0));
                whole = rayCast(c, rpos, rdir, hit, norm, clr, 1);
                if (hit.z == T(NOHIT)) {
                    for (int k = 0; k < 4; k++) res[k] = c.back[k];
                } else {
                    V3<T> ld = normalize(c.light - hit);
                    T diff = T(0.9) * std::max(T(0), dot(norm, ld));
                    T amb = T(0.1);
                    if (c.shadowamt > 0) {
                        V3<T> h2 = mk(T(0), T(0), T(NOHIT)), n2 = mk(T(0), T(0), T(0));
                        T c2[4] = {0, 0, 0, 1};
                        whole = rayCast(c, hit + norm * c.shadowbias, ld, h2, n2, c2, 1) && whole;
                        diff = (h2.z == T(NOHIT)) ? diff : diff * (1 - c.shadowamt);
                    }
                    for (int k = 0; k < 3; k++) res[k] = clr[k] * (diff + amb);
                    res[3] = 1;
                }
            }
            for (int k = 0; k < 4; k++) out[4 * (size_t(y) * W + x) + k] = float(res[k]);
            if (capped) capped[size_t(y) * W + x] = whole ? 0 : 1;
        }
}

// ============================================================ C ABI ====
// ------------------------------------------------------------- Compute ----
// Compute (src/gvdb_volume_gvdb.cpp:4527-4552) with the float operators of
// kernels/cuda_gvdb_operators.cuh: gvdbOpFillF (:371-384, every atlas voxel
// of the unpacked index range), gvdbOpSmooth (:400-417), gvdbOpNoise
// (:442-453), gvdbOpGrow (:235-253) (each over the interior of every used
// brick, reading the brick as it was before the launch).  The hash is
// random(float3) of kernels/cuda_gvdb.cuh:135-164 (Bob Jenkins'
// one-at-a-time final mix, compounded, low 23 bits as a float in [0, 1)).
static uint32_t oaat(uint32_t x) {
    x += (x << 10u);
    x ^= (x >> 6u);
    x += (x << 3u);
    x ^= (x >> 11u);
    x += (x << 15u);
    return x;
}
static float random3(float a, float b, float c) {
    uint32_t ia, ib, ic;
    memcpy(&ia, &a, 4); memcpy(&ib, &b, 4); memcpy(&ic, &c, 4);
    uint32_t m = (oaat(ia ^ oaat(ib) ^ oaat(ic)) & 0x007FFFFFu) | 0x3F800000u;
    float f;
    memcpy(&f, &m, 4);
    return f - 1.0f;
}

template <typename T>
static void computeT(Vol& v, int op, float p1, float p2, float p3, bool skip, std::vector<T>& atlas) {
    const int br = brickres(v);
    int res[3];
    atlasRes(v, res);
    if (op == 110) {
        int lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
            int count = skip ? v.atlas_cnt[k] * v.brickwid : res[k];
            int blocks = (count + 7) / 8;
            lo[k] = 1;
            hi[k] = std::min(8 * blocks, res[k] - 1);
        }
        for (int z = lo[2]; z <= hi[2]; z++)
            for (int y = lo[1]; y <= hi[1]; y++)
                for (int x = lo[0]; x <= hi[0]; x++) {
                    T val = T(p1);
                    if (p3 < 0) {
                        double ax = double(uint32_t(x) * 12u) / (3.141592 * 30.0);
                        double ay = double(uint32_t(y) * 12u) / (3.141592 * 30.0);
                        double az = double(uint32_t(z) * 12u) / (3.141592 * 30.0);
                        if (sizeof(T) == 4) {
                            float f = sinf(float(ax));
                            f += sinf(float(ay));
                            f += sinf(float(az));
                            val = T(f);
                        } else {
                            val = T(std::sin(double(float(ax))) + std::sin(double(float(ay))) + std::sin(double(float(az))));
                        }
                    }
                    atlas[aidx(v, x, y, z)] = val;
                }
        return;
    }
    const u64 slots = std::min<u64>(v.atlas_used, u64(v.atlas_cnt[0]) * v.atlas_cnt[1] * v.atlas_cnt[2]);
#pragma omp parallel for schedule(dynamic, 16)
    for (i64 s = 0; s < i64(slots); s++) {
        const int sx = int(s % v.atlas_cnt[0]), sy = int((s / v.atlas_cnt[0]) % v.atlas_cnt[1]);
        const int sz = int(s / (i64(v.atlas_cnt[0]) * v.atlas_cnt[1]));
        const int ox = sx * br, oy = sy * br, oz = sz * br;
        std::vector<T> b(size_t(br) * br * br);
        for (int z = 0; z < br; z++)
            for (int y = 0; y < br; y++)
                for (int x = 0; x < br; x++) b[(size_t(z) * br + y) * br + x] = atlas[aidx(v, ox + x, oy + y, oz + z)];
        auto at = [&](int x, int y, int z) { return b[(size_t(z) * br + y) * br + x]; };
        for (int z = 1; z < br - 1; z++)
            for (int y = 1; y < br - 1; y++)
                for (int x = 1; x < br - 1; x++) {
                    const T c = at(x, y, z);
                    T val = c;
                    if (op == 113) {
                        if (sizeof(T) == 4) {
                            float f = p1 * float(c);
                            f += float(at(x - 1, y, z));
                            f += float(at(x + 1, y, z));
                            f += float(at(x, y - 1, z));
                            f += float(at(x, y + 1, z));
                            f += float(at(x, y, z - 1));
                            f += float(at(x, y, z + 1));
                            val = T(float(double(f) / (double(p1) + 6.0) + double(p2)));
                        } else {
                            double d = double(p1) * double(c) + double(at(x - 1, y, z)) + double(at(x + 1, y, z)) +
                                       double(at(x, y - 1, z)) + double(at(x, y + 1, z)) + double(at(x, y, z - 1)) +
                                       double(at(x, y, z + 1));
                            val = T(d / (double(p1) + 6.0) + double(p2));
                        }
                    } else if (op == 114) {
                        if (double(c) > 0.01)
                            val = c + T(random3(float(ox + x), float(oy + y), float(oz + z))) * T(p1);
                    } else if (op == 115) {
                        if (c != T(0)) val = T(double(c) + double(p1) * 10.0);
                        if (double(val) < 0.01) val = T(0);
                    }
                    atlas[aidx(v, ox + x, oy + y, oz + z)] = val;
                }
    }
}

extern "C" {

void* orc_new(int levels, const int* logdims) {
    Vol* v = new Vol();
    int cnt[MAXLEV_HOST];
    // Configure(q4..q0) initial pool counts: src/gvdb_volume_gvdb.cpp:2364-2377
    int c = 4;
    cnt[0] = 4;
    for (int l = 1; l < levels; l++) { cnt[l] = c; c >>= 1; if (c < 1) c = 1; }
    Configure(*v, levels, logdims, cnt);
    return v;
}
void orc_free(void* h) { delete static_cast<Vol*>(h); }

void orc_set_channel_default(void* h, int x, int y, int z) {
    Vol& v = *static_cast<Vol*>(h);
    v.default_axiscnt[0] = x; v.default_axiscnt[1] = y; v.default_axiscnt[2] = z;
}

void orc_add_channel(void* h, int apron) {
    // AddChannel(0, T_FLOAT, apron): src/gvdb_volume_gvdb.cpp:2437-2453; AtlasCreate allocator.cpp:528-560
    Vol& v = *static_cast<Vol*>(h);
    v.has_atlas = true;
    v.apron = apron;
    v.brickwid = v.res[0];
    for (int a = 0; a < 3; a++) v.atlas_cnt[a] = v.default_axiscnt[a];
    v.atlas_max = u64(v.atlas_cnt[0]) * v.atlas_cnt[1] * v.atlas_cnt[2];
    int r[3];
    atlasRes(v, r);
    v.atlasF.assign(size_t(r[0]) * r[1] * r[2], 0.f);
    v.atlasD.assign(size_t(r[0]) * r[1] * r[2], 0.0);
}

// Keeps (1, the default) or drops (0) the FP64 shadow atlas; set before the
// first UpdateAtlas.  Without it only precision 32 runs.
void orc_set_fp64(void* h, int on) { static_cast<Vol*>(h)->fp64 = on != 0; }
void orc_set_epsilon(void* h, float eps) { static_cast<Vol*>(h)->epsilon = eps; }
int orc_undefined_events(void* h) { return static_cast<Vol*>(h)->undefined_events; }

int orc_rebuild_topology(void* h, const float* pts, int n, float radius, const float* orig) {
    Vol& v = *static_cast<Vol*>(h);
    auto t0 = std::chrono::steady_clock::now();
    int rc = rebuildTopology(v, pts, n, radius, orig);
    v.stage_ms[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int orc_accumulate_topology(void* h, const float* pts, int n, float radius, const float* orig, int iDepth) {
    Vol& v = *static_cast<Vol*>(h);
    return accumulateTopology(v, pts, n, radius, orig, iDepth);
}
void orc_request_full_rebuild(void* h, int on) {
    // RequestFullRebuild: src/gvdb_volume_gvdb.h:432
    static_cast<Vol*>(h)->rebuild_topo = on != 0;
}
unsigned long long orc_pool_used(void* h, int grp, int lev) { return static_cast<Vol*>(h)->pool[grp][lev].usedNum; }

int orc_rebuild_topology_cpu(void* h, const float* pts, int n, const float* orig) {
    // RebuildTopologyCPU: src/gvdb_volume_gvdb.cpp:1220-1227 (ActivateSpace per point, Vector3DI truncation)
    Vol& v = *static_cast<Vol*>(h);
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n; i++) {
        float p[3] = {pts[3 * i] + orig[0], pts[3 * i + 1] + orig[1], pts[3 * i + 2] + orig[2]};
        int ip[3] = {int(p[0]), int(p[1]), int(p[2])};
        bool bnew = false;
        ActivateSpaceR(v, v.root, ip, bnew, ID_UNDEFL, 0);
    }
    computeTopLev(v);
    v.stage_ms[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

void orc_clear(void* h) { Clear(*static_cast<Vol*>(h)); }

void orc_finish_topology(void* h) {
    Vol& v = *static_cast<Vol*>(h);
    computeBounds(v);
    computeTopLev(v);
}

int orc_update_atlas(void* h) {
    Vol& v = *static_cast<Vol*>(h);
    if (!v.has_atlas) return -3;
    auto t0 = std::chrono::steady_clock::now();
    updateAtlas(v);
    v.stage_ms[1] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

void orc_clear_channel(void* h) {
    Vol& v = *static_cast<Vol*>(h);
    std::fill(v.atlasF.begin(), v.atlasF.end(), 0.f);
    std::fill(v.atlasD.begin(), v.atlasD.end(), 0.0);
}

// Every atlas voxel (interior and apron) set to `value`: a uniform field for
// closed-form render checks (no reference counterpart; FillChannel-like).
void orc_fill_channel(void* h, float value) {
    Vol& v = *static_cast<Vol*>(h);
    std::fill(v.atlasF.begin(), v.atlasF.end(), value);
    std::fill(v.atlasD.begin(), v.atlasD.end(), double(value));
}

unsigned long long orc_pool_count(void* h, int grp, int lev) { return static_cast<Vol*>(h)->pool[grp][lev].lastEle; }
unsigned long long orc_pool_stride(void* h, int grp, int lev) { return static_cast<Vol*>(h)->pool[grp][lev].stride; }
void orc_pool_copy(void* h, int grp, int lev, void* dst) {
    Pool& p = static_cast<Vol*>(h)->pool[grp][lev];
    if (p.lastEle) memcpy(dst, p.data.data(), p.lastEle * p.stride);
}
int orc_top_lev(void* h) { return static_cast<Vol*>(h)->top_lev; }
void orc_bounds(void* h, float* out) {
    Vol& v = *static_cast<Vol*>(h);
    for (int a = 0; a < 3; a++) { out[a] = v.objMin[a]; out[3 + a] = v.objMax[a]; out[6 + a] = v.posMin[a]; out[9 + a] = v.posMax[a]; }
}
void orc_atlas_info(void* h, int* out) {
    Vol& v = *static_cast<Vol*>(h);
    int r[3];
    atlasRes(v, r);
    for (int a = 0; a < 3; a++) { out[a] = v.atlas_cnt[a]; out[3 + a] = r[a]; }
    out[6] = int(v.atlas_used);
}
void orc_atlas_copy(void* h, int precision, void* dst) {
    Vol& v = *static_cast<Vol*>(h);
    if (precision == 64) memcpy(dst, v.atlasD.data(), v.atlasD.size() * sizeof(double));
    else memcpy(dst, v.atlasF.data(), v.atlasF.size() * sizeof(float));
}
void orc_atlas_set(void* h, int precision, const void* src) {
    Vol& v = *static_cast<Vol*>(h);
    if (precision == 64) memcpy(v.atlasD.data(), src, v.atlasD.size() * sizeof(double));
    else memcpy(v.atlasF.data(), src, v.atlasF.size() * sizeof(float));
}
unsigned orc_hash(unsigned x) { return oaat(x); }
void orc_atlas_map_copy(void* h, void* dst) {
    Vol& v = *static_cast<Vol*>(h);
    memcpy(dst, v.amap.data(), v.amap.size() * sizeof(AtlasNode));
}

// the colour channel: AddChannel(chan, T_UCHAR4, apron) on the same atlas
void orc_add_color_channel(void* h) {
    Vol& v = *static_cast<Vol*>(h);
    v.has_clr = true;
    int r[3];
    atlasRes(v, r);
    v.atlasC.assign(size_t(r[0]) * r[1] * r[2], 0u);
}
void orc_set_color_channel(void* h, int on) { static_cast<Vol*>(h)->render_clr = on != 0; }
void orc_clear_color(void* h) {
    Vol& v = *static_cast<Vol*>(h);
    std::fill(v.atlasC.begin(), v.atlasC.end(), 0u);
}
void orc_color_copy(void* h, void* dst) {
    Vol& v = *static_cast<Vol*>(h);
    memcpy(dst, v.atlasC.data(), v.atlasC.size() * 4);
}
void orc_color_set(void* h, const void* src) {
    Vol& v = *static_cast<Vol*>(h);
    memcpy(v.atlasC.data(), src, v.atlasC.size() * 4);
}
// InsertPointsSubcell with SetPoints' colour buffer (a uint32 per point)
int orc_insert_points_subcell_clr(void* h, int scs, const float* pts, const uint32_t* clr, int n, float radius,
                                  const float* trans, int* outLen) {
    return insertPointsSubcell(*static_cast<Vol*>(h), scs, pts, n, radius, trans, outLen, 0, clr);
}
// GatherDensity's colour channel (the density part is orc_gather_density);
// exact = 0: the reference's float arithmetic in list order
int orc_gather_color(void* h, int scs, int accum, int exact) {
    Vol& v = *static_cast<Vol*>(h);
    if (!v.has_clr || v.sc_clr.size() != size_t(v.sc_len)) return -3;
    if (v.sc_len == 0) return 0;
    gatherColor(v, scs, accum != 0, exact != 0);
    return 0;
}
// UpdateApron<uchar4> (kernels/cuda_gvdb_operators.cuh:143-146): boundval
// converted to uchar per component
int orc_update_apron_color(void* h, float boundval) {
    Vol& v = *static_cast<Vol*>(h);
    if (!v.has_clr) return -3;
    if (v.apron == 0) return 0;
    const uint32_t b = uint32_t(toU8(boundval)) * 0x01010101u;
    apronT<uint32_t, float>(v, b, v.atlasC);
    return 0;
}

int orc_insert_points_subcell(void* h, int scs, const float* pts, int n, float radius, const float* trans, int* outLen) {
    Vol& v = *static_cast<Vol*>(h);
    auto t0 = std::chrono::steady_clock::now();
    int rc = insertPointsSubcell(v, scs, pts, n, radius, trans, outLen);
    v.stage_ms[2] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}
// InsertPointsSubcell_FP16 (gvdb_volume_gvdb.cpp:5094-5189): each entry
// coordinate u16((w - mPosMin) / mPosRange * 65535) (cuda_gvdb_particles.cuh:
// 204-207; float subtract, divide, multiply, truncation), against the bounds
// GetBoundingBox set (:5452-5462).  Outside [0, 65535] the reference's
// conversion is undefined; here it saturates.
static uint16_t quantize16(float w, float mn, float rg) {
    const float q = (w - mn) / rg * 65535.0f;
    if (q >= 65535.0f) return 65535;
    return q > 0.0f ? uint16_t(int(q)) : uint16_t(0);
}
int orc_insert_points_subcell_fp16(void* h, int scs, const float* pts, int n, float radius, const float* trans,
                                   int* outLen) {
    Vol& v = *static_cast<Vol*>(h);
    for (int a = 0; a < 3; a++) {
        v.qmin[a] = v.posMin[a];
        v.qrange[a] = v.posMax[a] - v.posMin[a];
        if (!(v.qrange[a] > 0.0f)) return -3;
    }
    int rc = insertPointsSubcell(v, scs, pts, n, radius, trans, outLen, 1);
    if (rc) return rc;
    v.sc_pnt16.resize(v.sc_pnt.size());
    for (size_t i = 0; i < v.sc_pnt.size(); i++) v.sc_pnt16[i] = quantize16(v.sc_pnt[i], v.qmin[i % 3], v.qrange[i % 3]);
    v.sc16 = true;
    return 0;
}
void orc_subcell16_copy(void* h, void* dst) {
    Vol& v = *static_cast<Vol*>(h);
    memcpy(dst, v.sc_pnt16.data(), v.sc_pnt16.size() * 2);
}
int orc_num_subcells(void* h) { return static_cast<Vol*>(h)->num_sc; }
void orc_subcell_copy(void* h, int which, void* dst) {
    Vol& v = *static_cast<Vol*>(h);
    switch (which) {
        case 0: memcpy(dst, v.sc_cnt.data(), v.sc_cnt.size() * 4); break;
        case 1: memcpy(dst, v.sc_off.data(), v.sc_off.size() * 4); break;
        case 2: memcpy(dst, v.sc_pos.data(), v.sc_pos.size() * 4); break;
        case 3: memcpy(dst, v.sc_nid.data(), v.sc_nid.size() * 4); break;
        case 4: memcpy(dst, v.sc_pnt.data(), v.sc_pnt.size() * 4); break;
        case 5: memcpy(dst, v.sc_mapping.data(), v.sc_mapping.size() * 4); break;
    }
}

int orc_gather_density(void* h, int scs, float radius, int precision, int accum) {
    Vol& v = *static_cast<Vol*>(h);
    if (precision == 64 && !v.fp64) return -1;
    if (v.sc_len == 0) return 0;
    auto t0 = std::chrono::steady_clock::now();
    if (precision == 64) gatherT<double>(v, scs, radius, accum != 0, v.atlasD);
    else gatherT<float>(v, scs, radius, accum != 0, v.atlasF);
    v.stage_ms[3] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

// GatherLevelSet (src/gvdb_volume_gvdb.cpp:5348-5371): no early return on an
// empty entry list, unlike GatherDensity.
int orc_gather_level_set(void* h, int scs, float radius, int precision, int accum) {
    Vol& v = *static_cast<Vol*>(h);
    if (precision == 64 && !v.fp64) return -1;
    if (v.pool[0][0].usedNum == 0) return 0;
    if (precision == 64) gatherT<double>(v, scs, radius, accum != 0, v.atlasD, true);
    else gatherT<float>(v, scs, radius, accum != 0, v.atlasF, true);
    return 0;
}

// GatherLevelSet_FP16 (src/gvdb_volume_gvdb.cpp:5374-5401): no accumulation.
int orc_gather_level_set_fp16(void* h, int scs, float radius, int precision) {
    Vol& v = *static_cast<Vol*>(h);
    if (precision == 64 && !v.fp64) return -1;
    if (v.pool[0][0].usedNum == 0) return 0;
    if (!v.sc16) return -3;
    if (precision == 64) gatherT<double>(v, scs, radius, false, v.atlasD, true, true);
    else gatherT<float>(v, scs, radius, false, v.atlasF, true, true);
    return 0;
}

int orc_update_apron(void* h, float boundval, int precision) {
    Vol& v = *static_cast<Vol*>(h);
    if (precision == 64 && !v.fp64) return -1;
    if (v.apron == 0) return 0;
    auto t0 = std::chrono::steady_clock::now();
    if (precision == 64) apronT<double>(v, double(boundval), v.atlasD);
    else apronT<float>(v, boundval, v.atlasF);
    v.stage_ms[4] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

float orc_random3(float a, float b, float c) { return random3(a, b, c); }

int orc_compute(void* h, int op, int iters, float p1, float p2, float p3, int update_apron, int skip, float boundval,
                int precision) {
    Vol& v = *static_cast<Vol*>(h);
    if (precision == 64 && !v.fp64) return -1;
    if (op != 110 && op != 113 && op != 114 && op != 115) return -7;
    for (int it = 0; it < iters; it++) {
        if (precision == 64) computeT<double>(v, op, p1, p2, p3, skip != 0, v.atlasD);
        else computeT<float>(v, op, p1, p2, p3, skip != 0, v.atlasF);
        if (update_apron && v.apron) {
            if (precision == 64) apronT<double>(v, double(boundval), v.atlasD);
            else apronT<float>(v, boundval, v.atlasF);
        }
    }
    return 0;
}

// Camera -> scene block (PrepareRender: src/gvdb_volume_gvdb.cpp:4254-4306).
// cam: angs[3], to[3], dist, dolly, fov, near, far (11 floats)
void orc_camera_scn(const float* cam, int w, int hgt, float* scn_out) {
    CamParams cp;
    memcpy(cp.angs, cam, 3 * 4);
    memcpy(cp.to, cam + 3, 3 * 4);
    cp.dist = cam[6]; cp.dolly = cam[7]; cp.fov = cam[8]; cp.znear = cam[9]; cp.zfar = cam[10];
    float pos[3], tl[3], tr[3], bl[3];
    cameraRays(cp, w, hgt, pos, tl, tr, bl);
    ScnBlock* s = reinterpret_cast<ScnBlock*>(scn_out);
    s->width = w; s->height = hgt;
    for (int a = 0; a < 3; a++) {
        s->campos[a] = pos[a];
        s->cams[a] = tl[a];
        s->camu[a] = tr[a] - tl[a];
        s->camv[a] = bl[a] - tl[a];
    }
}

int orc_scn_floats() { return int(sizeof(ScnBlock) / 4); }

// LinearTransferFunc: src/gvdb_scene.cpp:66-88 (on a zero-initialised LUT)
void orc_linear_tf(float t0, float t1, const float* a, const float* b, float* lut) {
    int n0 = int(t0 * float(TF_SIZE)), n1 = int(t1 * float(TF_SIZE));
    for (int n = n0; n < n1; n++) {
        float x = float(n - n0) / float(n1 - n0);
        for (int k = 0; k < 4; k++) lut[4 * n + k] = a[k] + x * (b[k] - a[k]);
    }
}

// capped (may be null): one byte per pixel, 1 where a ray stopped at MAX_ITER
int orc_render(void* h, const float* scn, const float* tf, int shading, int precision, float* out, uint8_t* capped) {
    Vol& v = *static_cast<Vol*>(h);
    if (precision == 64 && !v.fp64) return -1;
    const ScnBlock& s = *reinterpret_cast<const ScnBlock*>(scn);
    auto t0 = std::chrono::steady_clock::now();
    if (precision == 64) renderT<double>(v, v.atlasD, s, tf, shading, out, capped);
    else renderT<float>(v, v.atlasF, s, tf, shading, out, capped);
    v.stage_ms[5] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

double orc_stage_ms(void* h, int i) { return static_cast<Vol*>(h)->stage_ms[i]; }
int orc_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#endif
}

}  // extern "C"
